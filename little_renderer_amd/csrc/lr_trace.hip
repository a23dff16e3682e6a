// lr_trace.hip - closest-hit and any-hit traversal of ray batches through one binary BVH in
// the reference's GPU layout (32-byte nodes, 48-byte triangles) on gfx950: the walk of
// lr_walk.h, one ray per thread.
#include "lr_internal.h"
#include "lr_walk.h"

namespace {

constexpr int TBLOCK = 128;

// COUNT: also sums the node records fetched (the root's, an inner node's two children's, a
// popped node's) and the triangles tested, into counts[0], counts[1] (the algorithmic traffic)
template <bool ANY, bool COUNT>
__global__ __launch_bounds__(TBLOCK) void k_trace(const float4 *__restrict__ nodes, uint32_t root, const uint32_t *__restrict__ aux,
                                                    const float4 *__restrict__ tris, const float4 *__restrict__ rays, int n,
                                                    const uint32_t *count, float4 *__restrict__ hits, uint32_t *spill,
                                                    int spill_per, uint32_t *overflow, unsigned long long *counts)
{
    __shared__ uint32_t stk[LR_LDS_STACK][TBLOCK];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * TBLOCK + tid;
    const int live = count ? min(n, (int)*count) : n;   // a queue's count, on the device
    if (i >= live) return;
    const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
    const WalkStack S{&stk[0][tid], TBLOCK, spill + i, (size_t)n, LR_LDS_STACK + spill_per};
    uint32_t cn = 0, ct = 0;
    const WalkHit h = walk<ANY, false, COUNT>(nodes, root, aux, tris, nullptr, 0, v3of(r0), v3of(r1), r0.w, r1.w, S,
                                              overflow, cn, ct);
    hits[i] = make_float4(h.t, __uint_as_float(h.id), h.b1, h.b2);
    if (COUNT) {
        atomicAdd(&counts[0], (unsigned long long)cn);
        atomicAdd(&counts[1], (unsigned long long)ct);
    }
}

template <bool ANY, bool COUNT>
void launch(LrhContext *c, const uint32_t *nodes, uint32_t root, const uint32_t *aux, const float *tris, const float *rays,
            int nrays, const uint32_t *count, float *hits, int spill_per, unsigned long long *counts)
{
    hipLaunchKernelGGL((k_trace<ANY, COUNT>), dim3((nrays + TBLOCK - 1) / TBLOCK), dim3(TBLOCK), 0, c->stream,
                       reinterpret_cast<const float4 *>(nodes), root, aux, reinterpret_cast<const float4 *>(tris),
                       reinterpret_cast<const float4 *>(rays), nrays, count, reinterpret_cast<float4 *>(hits),
                       (uint32_t *)c->scratch[LR_SCRATCH_TRACE].base, spill_per, c->flags, counts);
}

}  // namespace

int lr_spill_per(int max_depth) { return max_depth + 2 > LR_LDS_STACK ? max_depth + 2 - LR_LDS_STACK : 0; }

int lr_launch_trace(LrhContext *c, const uint32_t *nodes, uint32_t root, const uint32_t *aux, const float *tris,
                    const float *rays, int nrays, const uint32_t *count, int any, float *hits, int max_depth,
                    unsigned long long *counts)
{
    if (nrays <= 0) return 0;
    if (max_depth < 1) return lr_fail(LRH_EINVAL, "lrh_trace: max_depth %d", max_depth);
    const int spill_per = lr_spill_per(max_depth);
    int rc = lr_scratch_reserve(c, LR_SCRATCH_TRACE, lr_align((size_t)spill_per * nrays * 4));
    if (rc) return rc;
    if (counts) {
        if (any) launch<true, true>(c, nodes, root, aux, tris, rays, nrays, count, hits, spill_per, counts);
        else launch<false, true>(c, nodes, root, aux, tris, rays, nrays, count, hits, spill_per, counts);
    } else {
        if (any) launch<true, false>(c, nodes, root, aux, tris, rays, nrays, count, hits, spill_per, nullptr);
        else launch<false, false>(c, nodes, root, aux, tris, rays, nrays, count, hits, spill_per, nullptr);
    }
    LR_HIP(hipGetLastError());
    return 0;
}
