// lr_photons.hip - the photon cache (include/lr_hip.h: lrh_photons_*): the reference viewer's
// caustics estimate between frames (gpu_trace_photons, engine/gpu/trace_gpu.c:1964-1996; deposit /
// pc_claim, trace.wgsl:1359-1404; its fading, trace_decay.wgsl; its lookup in the resolve,
// trace_resolve.wgsl:1-84). Light paths from the scene's lamp leave their power at every rough
// vertex after a mirror or glass in a world-space hash grid (the light kernels' photon form,
// lr_engine_k.h); the cache fades by a factor a frame; where the camera sees a rough surface
// directly it gives the caustics' radiance there: albedo / pi times the power over the area of
// the finest cells around the point that hold at least 16 photons. One difference: along the
// normal's main axis a point's cell is offset by 1/pi of a cell (pc_coord), so that a surface at
// a round coordinate (a floor at y = 0) does not lie on the boundary of two layers of cells.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "lr_engine_k.h"

extern "C" void lrs_brdf_tables(float *out);

namespace {

constexpr int PC_BLK = 256, PC_STK = 16;
constexpr uint32_t PC_CHUNK = 1u << 22;   // light paths a launch pair at most
constexpr size_t PC_SPILL_BYTES = (size_t)256 << 20;   // the walks' spill area at most (a chunk's share)
constexpr uint32_t PC_RESOLVE = 1u << 18;  // work-items of a lookup at most

__global__ __launch_bounds__(PC_BLK) void k_pc_clear(unsigned long long *t, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * PC_BLK + threadIdx.x; i < n; i += (size_t)gridDim.x * PC_BLK) t[i] = 0ull;
}

// the fade (trace_decay.wgsl): each held slot's sums and photons times k; a slot whose photons
// faded to nothing gives its slot back, but only when the next slot is free, so that a cell that
// probed past it is still found (a run of dead cells is given back from its end, a slot a fade)
__global__ __launch_bounds__(PC_BLK) void k_pc_fade(unsigned long long *t, double k)
{
    const uint32_t s = blockIdx.x * PC_BLK + threadIdx.x;
    if (s >= LR_PC_SLOTS) return;
    unsigned long long *w = t + 4 * (size_t)s;
    const uint64_t w3 = w[3];
    if ((uint32_t)(w3 >> 32) == 0u) return;   // free
    for (int ch = 0; ch < 3; ch++) w[ch] = (unsigned long long)llrint((double)w[ch] * k);
    const uint32_t count = (uint32_t)floor((double)(uint32_t)w3 * k);
    const unsigned long long next = __hip_atomic_load(&t[4 * (size_t)((s + 1u) & (LR_PC_SLOTS - 1u)) + 3], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
    if (count == 0u && (uint32_t)(next >> 32) == 0u) {
        w[0] = w[1] = w[2] = 0ull;
        __hip_atomic_store(&w[3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else w[3] = (w3 & 0xffffffff00000000ull) | count;
}

// a cell's rgb sums and photons (0 if it holds none), found from its home as pc_claim placed it
__device__ __forceinline__ float4 pc_cell_sum(const unsigned long long *t, uint64_t h)
{
    const uint32_t tag = pc_tag(h);
    uint32_t idx = (uint32_t)h & (LR_PC_SLOTS - 1u);
    for (int probe = 0; probe < 8; probe++) {
        const unsigned long long *w = t + 4 * (size_t)idx;
        const uint32_t k = (uint32_t)(w[3] >> 32);
        if (k == tag)
            return make_float4((float)w[0], (float)w[1], (float)w[2], (float)(uint32_t)w[3] / (float)LR_PC_UNIT);
        if (k == 0u) break;
        idx = (idx + 1u) & (LR_PC_SLOTS - 1u);
    }
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The lookup (trace_resolve.wgsl:43-80): each pixel's centre ray; where it first meets a rough
// surface, the cells around the point in the plane of its main axis, the finest size whose four
// nearest cells (bilinearly) hold at least 16 photons (else the largest): radiance = albedo / pi x
// their power x norm / size^2 (norm = 1 / (light paths x 256)). out: rgb, 1 (0 0 0 0 where the
// pixel sees no rough surface).
__global__ __launch_bounds__(PC_BLK) void k_pc_resolve(JobDev J, const unsigned long long *t, float norm, uint32_t *spill,
                                                       size_t spill_stride, int spill_per, const float *tab, uint32_t *overflow,
                                                       float4 *out)
{
    __shared__ uint32_t stk[PC_STK][PC_BLK];
    const size_t gtid = (size_t)blockIdx.x * PC_BLK + threadIdx.x;
    const WalkStack S{&stk[0][threadIdx.x], PC_BLK, spill + gtid, spill_stride, PC_STK + spill_per};
    for (size_t p = gtid; p < (size_t)J.npix; p += (size_t)gridDim.x * PC_BLK) {
        const int x = (int)(p % (uint32_t)J.W), y = (int)(p / (uint32_t)J.W);
        const float u = ((float)x + 0.5f) / (float)J.W, v = ((float)y + 0.5f) / (float)J.H;
        const V3 org = v3(J.cam[0], J.cam[1], J.cam[2]);
        const V3 ul = v3(J.cam[3], J.cam[4], J.cam[5]), hz = v3(J.cam[6], J.cam[7], J.cam[8]),
                 vt = v3(J.cam[9], J.cam[10], J.cam[11]);
        const V3 d = vnorm(vsub(vsub(vadd(ul, vscale(hz, u)), vscale(vt, v)), org));
        float4 res = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        WalkCnt wc;
        const WalkHit h = walk4<false, false, PC_STK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, org, d, J.eps, 1e30f, S,
                                                      overflow, wc);
        if (h.id != LR_NONE) {
            const V3 hp = vadd(org, vscale(d, h.t));
            V3 outward;
            int mi;
            uint32_t ii;
            hit_frame(J, org, d, h.t, h.id, __float_as_uint(h.b1), hp, outward, mi, ii);
            const bool front = vdot(d, outward) < 0.0f;
            const V3 nrm = front ? outward : vneg(outward);
            int type;
            const Bsdf b = bsdf_make(J.mats, mi, front, tab, type);
            if (type != MAT_LIGHT && (!b.smooth || (1.0f - b.metallic) * (1.0f - b.transmission) > 0.0f)) {
                const float4 alb = J.mats[3 * mi];
                const int axis = pc_axis(nrm), ax = axis / 2, t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
                const float pc3[3] = {hp.x, hp.y, hp.z};
                float size = J.pc_cell;
                float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                for (int level = 0; level < 3; level++) {
                    float fp[3];
                    int c0[3];
                    for (int k = 0; k < 3; k++) {
                        fp[k] = pc3[k] / size;
                        c0[k] = pc_coord(pc3[k], size, k == ax);
                    }
                    const float uu = fp[t1] - 0.5f, vv = fp[t2] - 0.5f;
                    const float u0 = floorf(uu), v0 = floorf(vv), fu = uu - u0, fv = vv - v0;
                    e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    for (int j = 0; j < 4; j++) {
                        int c[3] = {c0[0], c0[1], c0[2]};
                        c[t1] = (int)u0 + (j & 1);
                        c[t2] = (int)v0 + (j >> 1);
                        const float w = ((j & 1) ? fu : 1.0f - fu) * ((j >> 1) ? fv : 1.0f - fv);
                        const float4 cs = pc_cell_sum(t, pc_key(c[0], c[1], c[2], axis, level));
                        e.x += w * cs.x;
                        e.y += w * cs.y;
                        e.z += w * cs.z;
                        e.w += w * cs.w;
                    }
                    if (e.w >= 16.0f || level == 2) break;
                    size *= 4.0f;
                }
                const float k = norm / (size * size) / PI_F;
                res = make_float4(alb.x * e.x * k, alb.y * e.y * k, alb.z * e.z * k, 1.0f);
            }
        }
        out[p] = res;
    }
}

}  // namespace

struct LrhPhotons {
    LrhContext *ctx = nullptr;
    unsigned long long *table = nullptr;   // LR_PC_SLOTS x 4 words
    float cell = 0.0f;
    double paths = 0.0;                    // the light paths it holds, faded with it
    float *tab = nullptr;                  // the BSDF tables
    float4 *queue = nullptr;               // the light paths of a chunk that go on (queue_cap of them)
    uint32_t queue_cap = 0;
    uint32_t *qn = nullptr, *overflow = nullptr;
    unsigned long long *cnt = nullptr;     // the light kernels' ray counts (unused here)
    uint32_t *spill = nullptr;
    size_t spill_words = 0;
};

namespace {

// the walks' spill area for `threads` work-items over a tree `depth` deep
int pc_spill(LrhPhotons *pc, size_t threads, int depth, int *per)
{
    *per = depth + 2 - PC_STK > 0 ? depth + 2 - PC_STK : 0;
    const size_t words = threads * (size_t)(*per > 0 ? *per : 1);
    if (words <= pc->spill_words) return 0;
    LR_HIP(hipStreamSynchronize(pc->ctx->stream));
    if (pc->spill) (void)hipFree(pc->spill);
    pc->spill = nullptr;
    pc->spill_words = 0;
    if (hipMalloc(&pc->spill, words * 4) != hipSuccess) {
        (void)hipGetLastError();
        return lr_fail(LRH_ENOMEM, "photons: %zu bytes of walk stack could not be had", words * 4);
    }
    pc->spill_words = words;
    return 0;
}

// a chunk's light paths: at most PC_CHUNK, and as many as the spill area holds for the scene's tree;
// the queue made large enough for them
int pc_chunk(LrhPhotons *pc, int depth, uint32_t *chunk)
{
    const int per = depth + 2 - PC_STK > 0 ? depth + 2 - PC_STK : 1;
    const size_t by_spill = PC_SPILL_BYTES / ((size_t)per * 4) / PC_BLK * PC_BLK;
    *chunk = (uint32_t)std::max<size_t>(PC_BLK, std::min<size_t>(PC_CHUNK, by_spill));
    if (pc->queue_cap >= *chunk) return 0;
    LR_HIP(hipStreamSynchronize(pc->ctx->stream));
    if (pc->queue) (void)hipFree(pc->queue);
    pc->queue = nullptr;
    pc->queue_cap = 0;
    if (hipMalloc(&pc->queue, (size_t)*chunk * 64) != hipSuccess) {
        (void)hipGetLastError();
        return lr_fail(LRH_ENOMEM, "photons: the light paths' queue (%zu bytes) could not be had", (size_t)*chunk * 64);
    }
    pc->queue_cap = *chunk;
    return 0;
}

int pc_scene_check(const LrhPtScene *s)
{
    if (!s) return lr_fail(LRH_EINVAL, "photons: no scene");
    if (s->two_level) return lr_fail(LRH_EREFUSED, "photons: a scene of instances has no lamp");
    return 0;
}

}  // namespace

extern "C" {

int lrh_photons_create(LrhContext *c, float cell, LrhPhotons **out)
{
    if (!c || !out || !(cell > 0.0f)) return lr_fail(LRH_EINVAL, "lrh_photons_create: bad arguments");
    *out = nullptr;
    LrhPhotons *pc = new LrhPhotons;
    pc->ctx = c;
    pc->cell = cell;
    const size_t tbytes = (size_t)LR_PC_SLOTS * 32, qbytes = 0;
    char *m = nullptr;
    if (hipMalloc(&m, tbytes + qbytes + 2112 * 4 + 256) != hipSuccess) {
        (void)hipGetLastError();
        delete pc;
        return lr_fail(LRH_ENOMEM, "lrh_photons_create: %zu bytes of device memory could not be had", tbytes + qbytes);
    }
    pc->table = reinterpret_cast<unsigned long long *>(m);
    pc->tab = reinterpret_cast<float *>(m + tbytes + qbytes);
    pc->qn = reinterpret_cast<uint32_t *>(m + tbytes + qbytes + 2112 * 4);
    pc->overflow = pc->qn + 1;
    pc->cnt = reinterpret_cast<unsigned long long *>(pc->qn + 4);
    float tab[2112];
    lrs_brdf_tables(tab);
    if (hipMemcpy(pc->tab, tab, sizeof tab, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(pc->qn, 0, 256) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(m);
        delete pc;
        return lr_fail(LRH_EDEVICE, "lrh_photons_create: its tables could not be sent");
    }
    hipLaunchKernelGGL(k_pc_clear, dim3(4096), dim3(PC_BLK), 0, c->stream, pc->table, (size_t)LR_PC_SLOTS * 4);
    if (hipGetLastError() != hipSuccess) {
        (void)hipFree(m);
        delete pc;
        return lr_fail(LRH_EDEVICE, "lrh_photons_create: the table could not be cleared");
    }
    *out = pc;
    return 0;
}

void lrh_photons_destroy(LrhPhotons *pc)
{
    if (!pc) return;
    (void)hipStreamSynchronize(pc->ctx->stream);
    if (pc->spill) (void)hipFree(pc->spill);
    if (pc->queue) (void)hipFree(pc->queue);
    (void)hipFree(pc->table);
    delete pc;
}

int lrh_photons_clear(LrhPhotons *pc)
{
    if (!pc) return lr_fail(LRH_EINVAL, "lrh_photons_clear: no cache");
    hipLaunchKernelGGL(k_pc_clear, dim3(4096), dim3(PC_BLK), 0, pc->ctx->stream, pc->table, (size_t)LR_PC_SLOTS * 4);
    LR_HIP(hipGetLastError());
    pc->paths = 0.0;
    return 0;
}

int lrh_photons_trace(LrhPhotons *pc, const LrhPtScene *s, uint32_t npaths, uint32_t s0, uint64_t seed, int max_depth)
{
    if (!pc || npaths == 0 || max_depth < 0) return lr_fail(LRH_EINVAL, "lrh_photons_trace: bad arguments");
    int rc = pc_scene_check(s);
    if (rc) return rc;
    pc->paths += npaths;
    if (s->light < 0) return 0;   // no lamp: no light paths leave anything (they count all the same)
    uint32_t chunk = 0;
    rc = pc_chunk(pc, s->depth, &chunk);
    if (rc) return rc;
    int per = 0;
    rc = pc_spill(pc, chunk, s->depth, &per);
    if (rc) return rc;
    JobDev J;
    memset(&J, 0, sizeof(J));
    lr_job_scene(J, s);
    J.W = J.H = J.npix = 1;
    J.max_depth = max_depth;
    J.s0 = s0;
    J.seed_lo = (uint32_t)seed;
    J.seed_hi = (uint32_t)(seed >> 32);
    J.pc = pc->table;
    J.pc_cell = pc->cell;
    hipStream_t st = pc->ctx->stream;
    for (uint32_t g0 = 0; g0 < npaths; g0 += chunk) {
        const uint32_t n = std::min(chunk, npaths - g0);
        const unsigned blocks = (n + PC_BLK - 1) / PC_BLK;
        hipLaunchKernelGGL(k_light_reset, dim3(1), dim3(1), 0, st, pc->qn);
        hipLaunchKernelGGL((k_light_first<PC_BLK, PC_STK, false, true>), dim3(blocks), dim3(PC_BLK), 0, st, J, npaths,
                           (uint64_t)g0, n, pc->spill, (size_t)blocks * PC_BLK, per, pc->tab, pc->overflow, pc->queue, pc->qn,
                           pc->cnt);
        hipLaunchKernelGGL((k_light_rest<PC_BLK, PC_STK, false, true>), dim3(blocks), dim3(PC_BLK), 0, st, J, npaths, pc->spill,
                           (size_t)blocks * PC_BLK, per, pc->tab, pc->overflow, pc->queue, pc->qn, pc->cnt);
        LR_HIP(hipGetLastError());
    }
    return 0;
}

int lrh_photons_fade(LrhPhotons *pc, float k)
{
    if (!pc || !(k >= 0.0f && k <= 1.0f)) return lr_fail(LRH_EINVAL, "lrh_photons_fade: a factor in [0, 1]");
    hipLaunchKernelGGL(k_pc_fade, dim3(LR_PC_SLOTS / PC_BLK), dim3(PC_BLK), 0, pc->ctx->stream, pc->table, (double)k);
    LR_HIP(hipGetLastError());
    pc->paths *= k;
    return 0;
}

double lrh_photons_paths(const LrhPhotons *pc) { return pc ? pc->paths : 0.0; }

int lrh_photons_resolve(LrhPhotons *pc, const LrhPtScene *s, const float *cam, int W, int H, float *d_out)
{
    if (!pc || !cam || W <= 0 || H <= 0 || !d_out) return lr_fail(LRH_EINVAL, "lrh_photons_resolve: bad arguments");
    int rc = pc_scene_check(s);
    if (rc) return rc;
    const size_t npix = (size_t)W * (size_t)H;
    const size_t threads = std::min<size_t>((npix + PC_BLK - 1) / PC_BLK * PC_BLK, PC_RESOLVE);
    int per = 0;
    rc = pc_spill(pc, threads, s->depth, &per);
    if (rc) return rc;
    JobDev J;
    memset(&J, 0, sizeof(J));
    lr_job_scene(J, s);
    memcpy(J.cam, cam, sizeof(J.cam));
    J.W = W;
    J.H = H;
    J.npix = (int)npix;
    J.pc_cell = pc->cell;
    const float norm = pc->paths > 0.0 ? (float)(1.0 / (pc->paths * (double)LR_PC_SCALE)) : 0.0f;
    hipLaunchKernelGGL(k_pc_resolve, dim3((unsigned)(threads / PC_BLK)), dim3(PC_BLK), 0, pc->ctx->stream, J, pc->table, norm,
                       pc->spill, threads, per, pc->tab, pc->overflow, reinterpret_cast<float4 *>(d_out));
    LR_HIP(hipGetLastError());
    LR_HIP(hipStreamSynchronize(pc->ctx->stream));
    return 0;
}

int lrh_photons_read(LrhPhotons *pc, void *h_table)
{
    if (!pc || !h_table) return lr_fail(LRH_EINVAL, "lrh_photons_read: bad arguments");
    LR_HIP(hipStreamSynchronize(pc->ctx->stream));
    LR_HIP(hipMemcpy(h_table, pc->table, (size_t)LR_PC_SLOTS * 32, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
