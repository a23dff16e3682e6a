// lr_denoise.hip - the denoiser of the path tracer's images (include/lr_hip.h: lrh_denoise): an
// edge-avoiding a-trous wavelet filter (Dammertz et al. 2010) steered by each pixel's variance
// (SVGF, Schied et al. 2017, without its temporal part), the filter of the reference's
// engine/gpu/denoise_gpu.c and shaders/trace/denoise.wgsl:
//   prep    each pixel's irradiance (the mean of its samples over the guide's albedo, on a
//           surface) and the variance of that mean: from its own samples' moments once it has 4,
//           before that from the spread of its neighbours' (7 x 7, the same kind and facing)
//   passes  a 5 x 5 B3 spline with holes 1, 2, 4, 8, 16 pixels apart, each tap jittered within
//           its hole, weighted by luminance (against the noisier of the two pixels' variance),
//           normal (to the power 128) and distance (against its expected change); the variance
//           filtered alongside with the squared weights
//   last    times the albedo again; lamps and the sky pass as they are
// The passes are memory bound (25 taps, 5 times), so what they read is packed: 8 bytes of
// irradiance and variance (four halves) and 8 bytes of guide (the normal on an octahedron in
// two 16-bit snorms, the distance a float whose sign carries the kind), read in 16 x 16 tiles
// so a work-group's taps share the L2 lines of its neighbourhood.
#include <hip/hip_fp16.h>

#include "lr_internal.h"
#include "lr_pt_dev.h"

namespace {

constexpr int TILE = 16;   // a work-group: 16 x 16 pixels

// ---- packing (denoise.wgsl:12-45) ----
__device__ __forceinline__ uint32_t half2_bits(float a, float b)
{
    return (uint32_t)__half_as_ushort(__float2half_rn(a)) | ((uint32_t)__half_as_ushort(__float2half_rn(b)) << 16);
}
__device__ __forceinline__ float half_lo(uint32_t u) { return __half2float(__ushort_as_half((unsigned short)(u & 0xffffu))); }
__device__ __forceinline__ float half_hi(uint32_t u) { return __half2float(__ushort_as_half((unsigned short)(u >> 16))); }

// irradiance and variance (variance < 0: no sample yet)
__device__ __forceinline__ uint2 put_iv(V3 c, float v)
{
    return make_uint2(half2_bits(rfmin(c.x, 6e4f), rfmin(c.y, 6e4f)),
                      half2_bits(rfmin(c.z, 6e4f), rfmin(rfmax(v, -1.0f), 6e4f)));
}
__device__ __forceinline__ float4 get_iv(uint2 p) { return make_float4(half_lo(p.x), half_hi(p.x), half_lo(p.y), half_hi(p.y)); }

__device__ __forceinline__ uint32_t snorm16(float x)
{
    const float c = rfmin(rfmax(x, -1.0f), 1.0f);
    return (uint32_t)(uint16_t)(int16_t)floorf(0.5f + 32767.0f * c);
}
__device__ __forceinline__ float unsnorm16(uint32_t u) { return rfmax((float)(int16_t)(uint16_t)u / 32767.0f, -1.0f); }

// the normal on an octahedron (its lower half folded over)
__device__ __forceinline__ uint32_t oct_pack(V3 n)
{
    const float s = rfmax(fabsf(n.x) + fabsf(n.y) + fabsf(n.z), 1e-6f);
    float px = n.x / s, py = n.y / s;
    if (n.z < 0.0f) {
        const float fx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        const float fy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = fx;
        py = fy;
    }
    return snorm16(px) | (snorm16(py) << 16);
}
__device__ __forceinline__ V3 oct_unpack(uint32_t u)
{
    V3 n = v3(unsnorm16(u & 0xffffu), unsnorm16(u >> 16), 0.0f);
    n.z = 1.0f - fabsf(n.x) - fabsf(n.y);
    const float t = rfmax(-n.z, 0.0f);
    n.x += n.x >= 0.0f ? -t : t;
    n.y += n.y >= 0.0f ? -t : t;
    return vnorm(n);
}

struct Gd {
    V3 n;
    float z;
    int kind;   // 0 a surface, 1 a lamp or the sky
};
// the distance, and the kind in its sign: a surface z >= 0; a lamp or the sky -2e38
__device__ __forceinline__ uint2 put_g(V3 n, float z, float kind)
{
    const float e = kind == 0.0f ? z : -2e38f;
    return make_uint2(oct_pack(n), __float_as_uint(e));
}
__device__ __forceinline__ Gd get_g(uint2 p)
{
    Gd g;
    const float e = __uint_as_float(p.y);
    g.kind = e >= 0.0f ? 0 : 1;
    g.z = e >= 0.0f ? e : -e - 1.0f;
    g.n = oct_unpack(p.x);
    return g;
}

__device__ __forceinline__ V3 irradiance(float4 a, float4 alb)
{
    const float k = rfmax(a.w, 1.0f);
    const V3 mean = v3(a.x / k, a.y / k, a.z / k);
    return alb.w == 0.0f ? v3(mean.x / rfmax(alb.x, 1e-3f), mean.y / rfmax(alb.y, 1e-3f), mean.z / rfmax(alb.z, 1e-3f)) : mean;
}

// prep (denoise.wgsl:49-95): the packed guide, and each pixel's irradiance and its variance
__global__ __launch_bounds__(TILE *TILE) void k_dn_prep(const float4 *__restrict__ accum, const float2 *__restrict__ mom,
                                                        const float4 *__restrict__ guide, int W, int H, uint2 *__restrict__ gd,
                                                        uint2 *__restrict__ dst)
{
    const int x = blockIdx.x * TILE + (int)(threadIdx.x % TILE), y = blockIdx.y * TILE + (int)(threadIdx.x / TILE);
    if (x >= W || y >= H) return;
    const size_t k = (size_t)y * W + x;
    const float4 a = accum[k], gn = guide[2 * k], alb = guide[2 * k + 1];
    gd[k] = put_g(v3of(gn), gn.w, alb.w);
    if (a.w == 0.0f) {
        dst[k] = put_iv(v3(0.0f, 0.0f, 0.0f), -1.0f);
        return;
    }
    const V3 irr = irradiance(a, alb);
    if (alb.w != 0.0f) {
        dst[k] = put_iv(irr, 0.0f);
        return;
    }
    float v = 0.0f;
    if (a.w >= 4.0f) {
        const float2 m = mom[k];
        const float el = m.x / a.w;
        v = rfmax(m.y / a.w - el * el, 0.0f) / a.w;
    } else {
        // too few samples of its own: the spread of the neighbours' means, on the same surface
        float s1 = 0.0f, s2 = 0.0f, c = 0.0f;
        for (int dy = -3; dy <= 3; dy++)
            for (int dx = -3; dx <= 3; dx++) {
                const int qx = x + dx, qy = y + dy;
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                const size_t q = (size_t)qy * W + qx;
                const float4 aq = accum[q], albq = guide[2 * q + 1];
                if (aq.w == 0.0f || albq.w != alb.w) continue;
                if (vdot(v3of(guide[2 * q]), v3of(gn)) < 0.9f) continue;
                const float l = lum(irradiance(aq, albq));
                s1 += l;
                s2 += l * l;
                c += 1.0f;
            }
        const float el = s1 / c;
        v = rfmax(s2 / c - el * el, 0.0f);
    }
    dst[k] = put_iv(irr, v);
}

// one a-trous pass (denoise.wgsl:97-203) with holes `step` apart; the last writes the image
__global__ __launch_bounds__(TILE *TILE) void k_dn_atrous(const uint2 *__restrict__ src, const uint2 *__restrict__ gd,
                                                          const float4 *__restrict__ guide, int W, int H, int step, int last,
                                                          uint2 *__restrict__ dst, float4 *__restrict__ out)
{
    const int x = blockIdx.x * TILE + (int)(threadIdx.x % TILE), y = blockIdx.y * TILE + (int)(threadIdx.x / TILE);
    if (x >= W || y >= H) return;
    const size_t k = (size_t)y * W + x;
    const float4 c = get_iv(src[k]);
    const Gd gp = get_g(gd[k]);
    if (c.w < 0.0f || gp.kind == 1) {   // no sample yet; a lamp, the sky: as they are
        if (last) out[k] = c.w < 0.0f ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : make_float4(c.x, c.y, c.z, 1.0f);
        else dst[k] = src[k];
        return;
    }
    // the variance, blurred a little (3 x 3) so that one unlucky pixel does not stand out
    float v = 0.0f, vw = 0.0f;
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
            const int qx = min(max(x + dx, 0), W - 1), qy = min(max(y + dy, 0), H - 1);
            const float4 s = get_iv(src[(size_t)qy * W + qx]);
            if (s.w < 0.0f) continue;
            const float w = (dx != 0 ? 0.25f : 0.5f) * (dy != 0 ? 0.25f : 0.5f);
            v += w * s.w;
            vw += w;
        }
    v /= rfmax(vw, 1e-6f);
    const float lp = lum(v3(c.x, c.y, c.z));
    // how fast the distance changes from pixel to pixel here (the smaller side: an edge next to
    // the pixel does not count)
    const float zl = get_g(gd[(size_t)y * W + max(x - 1, 0)]).z, zr = get_g(gd[(size_t)y * W + min(x + 1, W - 1)]).z;
    const float zd = get_g(gd[(size_t)max(y - 1, 0) * W + x]).z, zu = get_g(gd[(size_t)min(y + 1, H - 1) * W + x]).z;
    const float gx = rfmin(fabsf(gp.z - zl), fabsf(gp.z - zr)), gy = rfmin(fabsf(gp.z - zd), fabsf(gp.z - zu));
    // the taps `step` apart, each moved at random within its hole (regular holes print a grid)
    uint32_t seed = hash32(((uint32_t)k * 0x9e3779b9u) ^ ((uint32_t)step * 0x85ebca6bu));
    const int jit = step / 2;
    const uint32_t jn = (uint32_t)(2 * jit + 1);
    const float hb[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    V3 sum = v3(0.0f, 0.0f, 0.0f);
    float vsum = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int j = 0; j < 5; j++)
#pragma unroll
        for (int i = 0; i < 5; i++) {
            int ox = (i - 2) * step, oy = (j - 2) * step;
            if (jit > 0 && (i != 2 || j != 2)) {
                seed = hash32(seed);
                ox += (int)(((seed & 0xffffu) * jn) >> 16) - jit;
                oy += (int)(((seed >> 16) * jn) >> 16) - jit;
            }
            const int qx = x + ox, qy = y + oy;
            if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
            const size_t q = (size_t)qy * W + qx;
            const float4 s = get_iv(src[q]);
            float w = hb[i] * hb[j];
            if (q != k) {
                const Gd gq = get_g(gd[q]);
                if (s.w < 0.0f || gq.kind != gp.kind) continue;
                // the noisier pixel's noise: a bright outlier spreads to its neighbours as much as
                // they take from it (the centre's alone would darken the image)
                w *= expf(-fabsf(lp - lum(v3(s.x, s.y, s.z))) / (4.0f * sqrtf(rfmax(v, s.w)) + 1e-6f));
                float wn = rfmax(vdot(gp.n, gq.n), 0.0f);
                for (int e = 0; e < 7; e++) wn *= wn;
                w *= wn;
                const float expect = gx * fabsf((float)ox) + gy * fabsf((float)oy);
                w *= expf(-fabsf(gp.z - gq.z) / (1.0f * expect + 1e-3f * gp.z + 1e-6f));
            }
            sum = vadd(sum, vscale(v3(s.x, s.y, s.z), w));
            vsum += w * w * s.w;
            wsum += w;
        }
    const V3 o = vscale(sum, 1.0f / wsum);
    if (last) {
        const float4 alb = guide[2 * k + 1];
        out[k] = make_float4(o.x * alb.x, o.y * alb.y, o.z * alb.z, 1.0f);
    } else
        dst[k] = put_iv(o, vsum / (wsum * wsum));
}

}  // namespace

size_t lr_denoise_scratch_bytes(int W, int H) { return 3 * lr_align((size_t)W * H * 8); }

int lr_denoise_launch(hipStream_t s, const float *accum, const float *moments, const float *guide, int W, int H, int passes,
                      void *scratch, float *out)
{
    const size_t n = (size_t)W * H;
    uint2 *a = (uint2 *)scratch, *b = (uint2 *)((char *)scratch + lr_align(n * 8)),
          *gd = (uint2 *)((char *)scratch + 2 * lr_align(n * 8));
    const dim3 grid((unsigned)((W + TILE - 1) / TILE), (unsigned)((H + TILE - 1) / TILE)), blk(TILE * TILE);
    hipLaunchKernelGGL(k_dn_prep, grid, blk, 0, s, (const float4 *)accum, (const float2 *)moments, (const float4 *)guide, W, H,
                       gd, a);
    // a -> b -> a ...; the last pass writes the image
    for (int i = 1; i <= passes; i++) {
        const uint2 *src = i % 2 ? a : b;
        uint2 *dst = i % 2 ? b : a;
        hipLaunchKernelGGL(k_dn_atrous, grid, blk, 0, s, src, (const uint2 *)gd, (const float4 *)guide, W, H, 1 << (i - 1),
                           i == passes ? 1 : 0, dst, (float4 *)out);
    }
    LR_HIP(hipGetLastError());
    return 0;
}

extern "C" int lrh_denoise(LrhContext *c, const float *d_accum, const float *d_moments, const float *d_guide, int w, int h,
                           int passes, float *d_out)
{
    if (!c || !d_accum || !d_moments || !d_guide || !d_out || w <= 0 || h <= 0 || passes < 1 || passes > 5)
        return lr_fail(LRH_EINVAL, "lrh_denoise: bad arguments (passes 1..5)");
    LR_HIP(hipSetDevice(c->device));
    int rc = lr_scratch_reserve(c, LR_SCRATCH_PT, lr_denoise_scratch_bytes(w, h));
    if (rc) return rc;
    return lr_denoise_launch(c->stream, d_accum, d_moments, d_guide, w, h, passes, c->scratch[LR_SCRATCH_PT].base, d_out);
}
