// lr_trace.hip - closest-hit and any-hit traversal of a binary BVH in the reference's GPU
// layout (32-byte nodes, 48-byte triangles) on gfx950.
//
// The walk and its arithmetic are trace.wgsl traverse (:270-392), hit_tri (:159-173) and
// box_t (:175-184) with safe_inv (media.wgsl:45-48): the same f32 operations in the same
// order (no fused multiply-add), so hits equal the oracle's to the bit on the same tree.
// The stack differs: the reference's 40 entries drop what does not fit; here the first
// LR_LDS_STACK entries of each ray live in LDS (interleaved by lane: no bank conflicts) and
// deeper ones spill to device memory, sized from the tree's depth, so nothing is dropped.
#include "lr_device.h"
#include "lr_internal.h"

namespace {

constexpr int TBLOCK = 128;
constexpr int LR_LDS_STACK = 24;

__device__ __forceinline__ float safe_inv(float d) { return fabsf(d) > 1e-12f ? 1.0f / d : 1e30f; }

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// box_t (trace.wgsl:175-184): the entry distance, -1 on a miss
__device__ __forceinline__ float box_t(float4 lo, float4 hi, V3 o, V3 inv, float tmin, float tmax)
{
    const float t0x = (lo.x - o.x) * inv.x, t0y = (lo.y - o.y) * inv.y, t0z = (lo.z - o.z) * inv.z;
    const float t1x = (hi.x - o.x) * inv.x, t1y = (hi.y - o.y) * inv.y, t1z = (hi.z - o.z) * inv.z;
    const float enter = wmax(wmax(wmin(t0x, t1x), wmin(t0y, t1y)), wmax(wmin(t0z, t1z), tmin));
    const float leave = wmin(wmin(wmax(t0x, t1x), wmax(t0y, t1y)), wmin(wmax(t0z, t1z), tmax));
    return leave < enter ? -1.0f : enter;
}

// hit_tri (trace.wgsl:159-173): t in (tmin, tmax) and the barycentrics
__device__ __forceinline__ bool hit_tri(float4 a, float4 b, float4 c, V3 o, V3 d, float tmin, float tmax, float &t_out,
                                        float &u_out, float &v_out)
{
    const V3 q = v3(a.x, a.y, a.z), u = v3(b.x, b.y, b.z), v = v3(c.x, c.y, c.z);
    const V3 pvec = vcross(d, v);
    const float det = vdot(u, pvec);
    if (fabsf(det) < 1e-12f) return false;
    const float inv = 1.0f / det;
    const V3 tvec = vsub(o, q);
    const float b1 = vdot(tvec, pvec) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    const V3 qvec = vcross(tvec, u);
    const float b2 = vdot(d, qvec) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    const float t = vdot(v, qvec) * inv;
    if (t <= tmin || t >= tmax) return false;
    t_out = t;
    u_out = b1;
    v_out = b2;
    return true;
}

// COUNT: also sums the node records fetched (the popped node, and an inner node's two
// children) and the triangles tested, into counts[0], counts[1] (the algorithmic traffic)
template <bool COUNT>
__global__ __launch_bounds__(TBLOCK) void k_trace(const float4 *__restrict__ nodes, uint32_t root, const uint32_t *__restrict__ aux,
                                                    const float4 *__restrict__ tris, const float4 *__restrict__ rays, int n,
                                                    const uint32_t *count, int any, float4 *__restrict__ hits, uint32_t *spill,
                                                    int spill_per, uint32_t *overflow, unsigned long long *counts)
{
    uint32_t cn = 0, ct = 0;
    __shared__ uint32_t stk[LR_LDS_STACK][TBLOCK];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * TBLOCK + tid;
    const int live = count ? min(n, (int)*count) : n;   // a queue's count, on the device
    if (i >= live) return;
    const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
    const V3 o = v3(r0.x, r0.y, r0.z), d = v3(r1.x, r1.y, r1.z);
    const float tmin = r0.w;
    const V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    float best_t = r1.w, bu = 0.0f, bv = 0.0f;
    uint32_t best_id = LR_NONE;
    const int cap = LR_LDS_STACK + spill_per;
    int sp = 0;
    auto push = [&](uint32_t x) {
        if (sp < LR_LDS_STACK) stk[sp][tid] = x;
        else spill[(size_t)(sp - LR_LDS_STACK) * n + i] = x;
        sp++;
    };
    push(root);
    while (sp > 0) {
        sp--;
        const uint32_t ni = sp < LR_LDS_STACK ? stk[sp][tid] : spill[(size_t)(sp - LR_LDS_STACK) * n + i];
        const float4 nlo = nodes[2 * (size_t)ni], nhi = nodes[2 * (size_t)ni + 1];
        const uint32_t na = __float_as_uint(nlo.w), nb = __float_as_uint(nhi.w);
        if (COUNT) cn++;
        if (nb & LR_LEAF) {
            const uint32_t cnt = nb & 0x7fffffffu;
            bool stop = false;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t sid = aux[na + k];
                const float4 ta = tris[3 * (size_t)sid], tb = tris[3 * (size_t)sid + 1], tc = tris[3 * (size_t)sid + 2];
                float t, u, v;
                if (COUNT) ct++;
                if (hit_tri(ta, tb, tc, o, d, tmin, best_t, t, u, v)) {
                    best_t = t;
                    best_id = sid;
                    bu = u;
                    bv = v;
                    if (any) { stop = true; break; }
                }
            }
            if (stop) break;
            continue;
        }
        const float4 alo = nodes[2 * (size_t)na], ahi = nodes[2 * (size_t)na + 1];
        const float4 blo = nodes[2 * (size_t)nb], bhi = nodes[2 * (size_t)nb + 1];
        const float tl = box_t(alo, ahi, o, inv, tmin, best_t);
        const float tr = box_t(blo, bhi, o, inv, tmin, best_t);
        if (COUNT) cn += 2;
        if (sp + 2 > cap) {   // cannot happen with a correct depth: said, never silently dropped
            atomicOr(overflow, 1u);
            break;
        }
        if (tl >= 0.0f && tr >= 0.0f) {
            const bool left_first = tl <= tr;
            push(left_first ? nb : na);
            push(left_first ? na : nb);
        } else if (tl >= 0.0f) push(na);
        else if (tr >= 0.0f) push(nb);
    }
    hits[i] = make_float4(best_t, __uint_as_float(best_id), bu, bv);
    if (COUNT) {
        atomicAdd(&counts[0], (unsigned long long)cn);
        atomicAdd(&counts[1], (unsigned long long)ct);
    }
}

}  // namespace

int lr_launch_trace(LrhContext *c, const uint32_t *nodes, uint32_t root, const uint32_t *aux, const float *tris,
                    const float *rays, int nrays, const uint32_t *count, int any, float *hits, int max_depth,
                    unsigned long long *counts)
{
    if (nrays <= 0) return 0;
    if (max_depth < 1) return lr_fail(LRH_EINVAL, "lrh_trace: max_depth %d", max_depth);
    const int need = max_depth + 2;
    const int spill_per = need > LR_LDS_STACK ? need - LR_LDS_STACK : 0;
    const size_t spill_bytes = lr_align((size_t)spill_per * nrays * 4);
    int rc = lr_scratch_reserve(c, LR_SCRATCH_TRACE, spill_bytes);
    if (rc) return rc;
    if (counts)
        hipLaunchKernelGGL(k_trace<true>, dim3((nrays + TBLOCK - 1) / TBLOCK), dim3(TBLOCK), 0, c->stream,
                           reinterpret_cast<const float4 *>(nodes), root, aux, reinterpret_cast<const float4 *>(tris),
                           reinterpret_cast<const float4 *>(rays), nrays, count, any, reinterpret_cast<float4 *>(hits),
                           (uint32_t *)c->scratch[LR_SCRATCH_TRACE].base, spill_per, c->flags, counts);
    else
        hipLaunchKernelGGL(k_trace<false>, dim3((nrays + TBLOCK - 1) / TBLOCK), dim3(TBLOCK), 0, c->stream,
                           reinterpret_cast<const float4 *>(nodes), root, aux, reinterpret_cast<const float4 *>(tris),
                           reinterpret_cast<const float4 *>(rays), nrays, count, any, reinterpret_cast<float4 *>(hits),
                           (uint32_t *)c->scratch[LR_SCRATCH_TRACE].base, spill_per, c->flags, nullptr);
    LR_HIP(hipGetLastError());
    return 0;
}
