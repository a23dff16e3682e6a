// lr_pt.hip - the wavefront path tracer on gfx950: the light transport of
// engine/rt/tracer.c (trace_path_sampled :364-509, trace_image :329-356) and of the reference
// GPU's wavefront kernels (trace_wave.wgsl: wf_gen, wf_ext, wf_shade, wf_shadow, wf_resolve),
// with the sampler of engine/base/sampler.c and the BSDF of engine/scene/brdf.h.
//
// A pass traces P = W x H x S paths (S samples of every pixel; path p is pixel p mod W*H,
// sample s0 + p / (W*H)). Per bounce: the extension kernel walks the ray queue (closest hit);
// the shade kernel adds emission, makes the next-event shadow ray (its contribution waiting
// in the shadow queue), samples the BSDF and appends the next ray; the shadow kernel walks the
// shadow queue (any hit) and adds the contributions that are not blocked. Queues are
// compacted per wave: one atomic per wave (ballot), each lane's slot by a popcount of the
// lanes below it. Every random number derives from (seed, pixel, sample, dimension) only: the
// sampler's 13 pairs, then a PCG32 stream per (pixel, sample), so the image is the same
// whatever the order, the batch or the device it is traced in.
#include <string.h>

#include <vector>

#include "lr_internal.h"
#include "lr_walk.h"

namespace {

constexpr int PBLOCK = 128;
constexpr float PI_F = 3.14159265358979323846f;
constexpr int PAIRS = 13;                  // SAMPLER_PAIRS (sampler.h:29)
enum { EMIT_FULL = 0, EMIT_NONE = 1, EMIT_MIS = 2 };
enum { MAT_DIFFUSE = 0, MAT_MIRROR = 1, MAT_GLASS = 2, MAT_LIGHT = 3, MAT_PBR = 4 };
enum { SPEC = 0, DIFF = 1, TRANS = 2, MS = 3, LOBES = 4 };

struct PtScene {
    const float4 *nodes;
    const uint32_t *aux;
    const float4 *tris;
    const float4 *prims;
    const float4 *mats;      // 3 float4 per material: albedo.rgb type | emission.rgb ior | metallic roughness transmission 0
    const float *tab;        // E[32][32] B[32][32] E_avg[32] B_avg[32]
    const uint16_t *mask;    // blue-noise ranks, 64 x 64
    uint32_t root, nprims;
    int light;
    float eps, margin;
};

struct PtJob {
    float cam[12];
    int W, H, S;             // image, samples per pixel in this pass
    uint32_t s0;             // the pass's first sample index
    uint32_t seed_lo, seed_hi;
    int max_depth, mis;
    float clamp;
};

struct PtBufs {
    float4 *thr;             // throughput.rgb, bounce_pdf
    float4 *rad;             // radiance.rgb, emit state
    uint64_t *rng;           // PCG32 state (its stream: the pixel)
    float4 *qray[2];         // o.xyz tmin | d.xyz tmax
    uint32_t *qpath[2];
    float4 *hits;            // t, id, b1, b2
    float4 *sray;
    uint32_t *spath;
    float4 *scon;            // the contribution a shadow ray adds when nothing blocks it
    uint32_t *count;         // [0], [1] the ray queues, [2] the shadow queue
    uint32_t *spill;
    int spill_per;
    uint32_t *overflow;
};

// ---- random numbers: PCG32 (rng.h:13-37) ----
__device__ __forceinline__ uint32_t pcg_next(uint64_t &state, uint64_t inc)
{
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
__device__ __forceinline__ float pcg_float(uint64_t &state, uint64_t inc)
{
    return (float)(pcg_next(state, inc) >> 8) * (1.0f / 16777216.0f);
}
// rng_seed (rng.h:22-28) of the stream for (pixel, sample k): the pixel's seed of
// tracer.c:339, xor the sample index times a constant; stream = the pixel
__device__ __forceinline__ uint64_t pcg_seed(uint64_t pixel, uint32_t k, uint64_t seed)
{
    const uint64_t inc = (pixel << 1u) | 1u;
    uint64_t st = 0;
    pcg_next(st, inc);
    st += 0x853c49e6748fea9bULL ^ (pixel * 0x9E3779B97F4A7C15ULL) ^ (seed * 0xD1B54A32D192ED03ULL) ^
          ((uint64_t)k * 0xA0761D6478BD642FULL);
    pcg_next(st, inc);
    return st;
}

// ---- the sampler (sampler.c:7-175; trace.wgsl:81-126) ----
__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t owen_r(uint32_t x, uint32_t seed)
{
    x ^= x * 0x3d20adeau;
    x += seed;
    x *= (seed >> 16) | 1u;
    x ^= x * 0x05526c56u;
    x ^= x * 0x53a22864u;
    return __brev(x);
}
__device__ __forceinline__ uint32_t sobol1_r(uint32_t i)
{
    i ^= (i >> 1) & 0x55555555u;
    i ^= (i >> 2) & 0x33333333u;
    i ^= (i >> 4) & 0x0f0f0f0fu;
    i ^= (i >> 8) & 0x00ff00ffu;
    return i ^ (i >> 16);
}
struct Smp {
    int x, y;
    uint32_t index, seed;
    const uint16_t *mask;
    uint64_t *st;
    uint64_t inc;
};
__device__ __forceinline__ float bn_value(const uint16_t *m, int x, int y)
{
    return ((float)m[(y & 63) * 64 + (x & 63)] + 0.5f) * (1.0f / 4096.0f);
}
__device__ __forceinline__ float wrap01(float t) { return t >= 1.0f ? t - 1.0f : t; }
// sampler_2d: pair `dim` (Owen-scrambled Sobol, its order shuffled, shifted by the blue
// noise), or two plain numbers past the sequence's pairs
__device__ __forceinline__ void s2(const Smp &s, int dim, float &a, float &b)
{
    if (dim >= PAIRS) {
        a = pcg_float(*s.st, s.inc);
        b = pcg_float(*s.st, s.inc);
        return;
    }
    const uint32_t tile = hash32(((uint32_t)(s.x >> 6) * 73856093u) ^ ((uint32_t)(s.y >> 6) * 19349663u));
    const uint32_t seed = hash32(s.seed ^ hash32(((uint32_t)dim * 0x9e3779b9u) ^ tile));
    const uint32_t i = owen_r(__brev(s.index), seed);
    const uint32_t xa = owen_r(i, hash32(seed ^ 0xa511e9b3u));
    const uint32_t xb = owen_r(sobol1_r(i), hash32(seed ^ 0x63d83595u));
    const int ox = s.x + 23 * dim, oy = s.y + 41 * dim;
    a = wrap01((float)(xa >> 8) * (1.0f / 16777216.0f) + bn_value(s.mask, ox, oy));
    b = wrap01((float)(xb >> 8) * (1.0f / 16777216.0f) + bn_value(s.mask, ox + 32, oy + 17));
}

// ---- vector helpers ----
__device__ __forceinline__ float vlen(V3 a) { return sqrtf(vdot(a, a)); }
__device__ __forceinline__ V3 vnorm(V3 a)
{
    const float l = vlen(a);
    return l > 0.0f ? vscale(a, 1.0f / l) : a;
}
__device__ __forceinline__ float vmax3(V3 a) { return rfmax(a.x, rfmax(a.y, a.z)); }
__device__ __forceinline__ V3 vreflect(V3 d, V3 n) { return vsub(d, vscale(n, 2.0f * vdot(d, n))); }
__device__ __forceinline__ V3 vrefract(V3 d, V3 n, float eta)
{
    const float cos_i = rfmin(vdot(vneg(d), n), 1.0f);
    const V3 perp = vscale(vadd(d, vscale(n, cos_i)), eta);
    const V3 para = vscale(n, -sqrtf(fabsf(1.0f - vdot(perp, perp))));
    return vadd(perp, para);
}
__device__ __forceinline__ void onb(V3 n, V3 &t, V3 &b)            // rng.h:53-59 (Duff et al. 2017)
{
    const float s = copysignf(1.0f, n.z);
    const float a = -1.0f / (s + n.z);
    const float c = n.x * n.y * a;
    t = v3(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    b = v3(c, s + n.y * n.y * a, -n.y);
}
__device__ __forceinline__ V3 cosine_hemisphere(V3 n, float u1, float u2)   // rng.h:63-70
{
    const float phi = 2.0f * PI_F * u1;
    const float rad = sqrtf(u2);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float x = rad * cp, y = rad * sp, z = sqrtf(rfmax(0.0f, 1.0f - u2));
    V3 t, b;
    onb(n, t, b);
    return vadd(vadd(vscale(t, x), vscale(b, y)), vscale(n, z));
}

// ---- the BSDF (brdf.h:43-348) ----
struct Bsdf {
    V3 base;
    float metallic, transmission, rough, alpha, eta, e_avg, es_avg;
    bool smooth;
    V3 ms;
    float kd_norm, kms_norm;
    float cv, p[LOBES], es_v, e_v;
};
__device__ __forceinline__ float clamp01(float x) { return x < 0.0f ? 0.0f : x > 1.0f ? 1.0f : x; }
__device__ __forceinline__ float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
__device__ __forceinline__ float table(const float *t, float rough, float mu)
{
    const float x = rough * 31.0f, y = clamp01(mu) * 31.0f;
    int j = (int)x, i = (int)y;
    if (j > 30) j = 30;
    if (i > 30) i = 30;
    const float fx = x - (float)j, fy = y - (float)i;
    const float e0 = t[j * 32 + i] + fy * (t[j * 32 + i + 1] - t[j * 32 + i]);
    const float e1 = t[(j + 1) * 32 + i] + fy * (t[(j + 1) * 32 + i + 1] - t[(j + 1) * 32 + i]);
    return e0 + fx * (e1 - e0);
}
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta)
{
    cos_i = clamp01(cos_i);
    const float s2 = eta * eta * (1.0f - cos_i * cos_i);
    if (s2 >= 1.0f) return 1.0f;
    const float c = eta > 1.0f ? sqrtf(1.0f - s2) : cos_i;
    float f0 = (1.0f - eta) / (1.0f + eta);
    f0 *= f0;
    const float x = 1.0f - c, x2 = x * x;
    return f0 + (1.0f - f0) * x2 * x2 * x;
}
__device__ __forceinline__ float fresnel_dielectric_avg(float eta)
{
    const float e = eta > 1.0f ? 1.0f / eta : eta;
    float f0 = (1.0f - e) / (1.0f + e);
    f0 *= f0;
    const float avg = f0 + (1.0f - f0) / 21.0f;
    return eta > 1.0f ? 1.0f - (1.0f - avg) / (eta * eta) : avg;
}
__device__ __forceinline__ float spec_albedo(const Bsdf &b, const float *tab, float mu)
{
    if (b.smooth || b.eta > 1.0f) return fresnel_dielectric(mu, b.eta);
    float f0 = (1.0f - b.eta) / (1.0f + b.eta);
    const float e = table(tab, b.rough, mu), bb = table(tab + 1024, b.rough, mu);
    f0 *= f0;
    return f0 * (e - bb) + bb;
}
__device__ __forceinline__ V3 fresnel_metal(V3 f0, float c)
{
    const float x = 1.0f - clamp01(c), x2 = x * x, x5 = x2 * x2 * x;
    return v3(f0.x + (1.0f - f0.x) * x5, f0.y + (1.0f - f0.y) * x5, f0.z + (1.0f - f0.z) * x5);
}
__device__ __forceinline__ float ggx_d(float nh, float s2, float a)
{
    const float a2 = a * a, d = s2 + a2 * nh * nh;
    return a2 / (PI_F * d * d);
}
__device__ __forceinline__ float ggx_s2(V3 n, V3 h)
{
    const V3 c = vcross(n, h);
    return vdot(c, c);
}
__device__ __forceinline__ float ggx_g1(float c, float a)
{
    c = fabsf(c);
    return 2.0f * c / (c + sqrtf(a * a + (1.0f - a * a) * c * c));
}
__device__ __forceinline__ float ggx_vis(float cv, float cl, float a)
{
    cv = fabsf(cv);
    cl = fabsf(cl);
    const float a2 = a * a;
    const float d = cl * sqrtf(cv * cv * (1.0f - a2) + a2) + cv * sqrtf(cl * cl * (1.0f - a2) + a2);
    return d > 0.0f ? 0.5f / d : 0.0f;
}
__device__ __forceinline__ V3 ggx_sample_vndf(V3 n, V3 v, float a, float u1, float u2)   // Dupuy & Benyoub 2023
{
    V3 t, bt;
    onb(n, t, bt);
    const V3 vh = vnorm(v3(a * vdot(v, t), a * vdot(v, bt), vdot(v, n)));
    const float phi = 2.0f * PI_F * u1;
    const float z = (1.0f - u2) * (1.0f + vh.z) - vh.z;
    const float st = sqrtf(rfmax(0.0f, 1.0f - z * z));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const V3 hs = v3(st * cp + vh.x, st * sp + vh.y, z + vh.z);
    const V3 hl = vnorm(v3(a * hs.x, a * hs.y, rfmax(hs.z, 0.0f)));
    return vadd(vadd(vscale(t, hl.x), vscale(bt, hl.y)), vscale(n, hl.z));
}

// material_resolved (scene.h:80-90) + bsdf_make (brdf.h:176-205)
__device__ Bsdf bsdf_make(const float4 *mats, int mi, bool outside, const float *tab, int &type)
{
    const float4 m0 = mats[3 * mi], m1 = mats[3 * mi + 1], m2 = mats[3 * mi + 2];
    type = (int)m0.w;
    float metallic = m2.x, roughness = m2.y, transmission = m2.z, ior = m1.w;
    switch (type) {
    case MAT_DIFFUSE: metallic = 0.0f; roughness = 1.0f; ior = 1.5f; transmission = 0.0f; break;
    case MAT_MIRROR: metallic = 1.0f; roughness = 0.0f; ior = 1.5f; transmission = 0.0f; break;
    case MAT_GLASS: metallic = 0.0f; roughness = 0.0f; transmission = 1.0f; if (!(ior > 0.0f)) ior = 1.5f; break;
    case MAT_LIGHT: metallic = 0.0f; roughness = 1.0f; transmission = 0.0f; break;
    default: if (!(ior > 0.0f)) ior = 1.5f; break;
    }
    Bsdf b;
    b.base = v3(m0.x, m0.y, m0.z);
    b.metallic = clamp01(metallic);
    b.transmission = clamp01(transmission);
    const float r = clamp01(roughness);
    b.smooth = r <= 0.02f;
    b.rough = r;
    b.alpha = rfmax(r * r, 1e-4f);
    ior = ior > 0.0f ? ior : 1.5f;
    b.eta = b.transmission > 0.0f && !outside ? ior : 1.0f / ior;
    const float x = r * 31.0f;
    int j = (int)x;
    if (j > 30) j = 30;
    const float *EA = tab + 2048, *BA = tab + 2080;
    b.e_avg = EA[j] + (x - (float)j) * (EA[j + 1] - EA[j]);
    const float b_avg = BA[j] + (x - (float)j) * (BA[j + 1] - BA[j]);
    float f0 = (1.0f - b.eta) / (1.0f + b.eta);
    f0 *= f0;
    b.es_avg = f0 * (b.e_avg - b_avg) + b_avg;
    const float sa = b.smooth || b.eta > 1.0f ? fresnel_dielectric_avg(b.eta) : b.es_avg;
    b.kd_norm = 1.0f / (1.0f - sa);
    b.kms_norm = 1.0f / (PI_F * (1.0f - b.e_avg));
    if (b.smooth || b.metallic <= 0.0f) b.ms = v3(0.0f, 0.0f, 0.0f);
    else {
        const float e = b.e_avg;
        const V3 fa = v3(b.base.x + (1.0f - b.base.x) / 21.0f, b.base.y + (1.0f - b.base.y) / 21.0f,
                         b.base.z + (1.0f - b.base.z) / 21.0f);
        b.ms = v3(fa.x * fa.x * e / (1.0f - fa.x * (1.0f - e)), fa.y * fa.y * e / (1.0f - fa.y * (1.0f - e)),
                  fa.z * fa.z * e / (1.0f - fa.z * (1.0f - e)));
    }
    b.cv = -1.0f;
    b.es_v = b.e_v = 0.0f;
    for (int i = 0; i < LOBES; i++) b.p[i] = 0.0f;
    return b;
}

__device__ void bsdf_view(Bsdf &b, const float *tab, float cv)   // brdf.h:209-224
{
    if (b.cv == cv) return;
    b.cv = cv;
    const float fd = fresnel_dielectric(cv, b.eta);
    const float fs = b.smooth ? fd : rfmin(fd, 0.9f);
    const float m = b.metallic, t = b.transmission, lb = lum(b.base);
    b.es_v = (1.0f - m) * (1.0f - t) > 0.0f ? spec_albedo(b, tab, cv) : 0.0f;
    b.e_v = b.smooth || m <= 0.0f ? 1.0f : table(tab, b.rough, cv);
    float w[LOBES];
    w[SPEC] = (1.0f - m) * fs + m * lum(fresnel_metal(b.base, cv));
    w[DIFF] = (1.0f - m) * (1.0f - t) * (1.0f - b.es_v) * lb;
    w[TRANS] = (1.0f - m) * t * (1.0f - fs) * lb;
    w[MS] = m * lum(b.ms) * (1.0f - b.e_v);
    const float sum = w[0] + w[1] + w[2] + w[3];
    for (int k = 0; k < LOBES; k++) b.p[k] = sum > 0.0f ? w[k] / sum : 0.0f;
}

__device__ V3 bsdf_eval(Bsdf &b, const float *tab, V3 n, V3 v, V3 l, float &pdf)   // brdf.h:228-275
{
    V3 f = v3(0.0f, 0.0f, 0.0f);
    pdf = 0.0f;
    const float cv = vdot(v, n), cl = vdot(l, n);
    if (cv <= 0.0f) return f;
    bsdf_view(b, tab, cv);
    const float m = b.metallic, t = b.transmission, a = b.alpha;
    if (cl > 0.0f) {
        const V3 h = vnorm(vadd(v, l));
        const float vh = clamp01(vdot(v, h)), nh = vdot(n, h);
        const float fd = fresnel_dielectric(vh, b.eta);
        float kd = (1.0f - m) * (1.0f - t);
        if (kd > 0.0f) {
            kd *= (1.0f - b.es_v) * (1.0f - spec_albedo(b, tab, cl)) * b.kd_norm;
            f = vscale(b.base, kd / PI_F);
            pdf += b.p[DIFF] * cl / PI_F;
        }
        if (!b.smooth && m > 0.0f) {
            const float k = m * (1.0f - b.e_v) * (1.0f - table(tab, b.rough, cl)) * b.kms_norm;
            f = vadd(f, vscale(b.ms, k));
            pdf += b.p[MS] * cl / PI_F;
        }
        if (!b.smooth && nh > 0.0f) {
            const float d = ggx_d(nh, ggx_s2(n, h), a);
            const V3 fm = fresnel_metal(b.base, vh);
            const float s = d * ggx_vis(cv, cl, a);
            f = vadd(f, v3(((1.0f - m) * fd + m * fm.x) * s, ((1.0f - m) * fd + m * fm.y) * s, ((1.0f - m) * fd + m * fm.z) * s));
            pdf += b.p[SPEC] * ggx_g1(cv, a) * d / (4.0f * cv);
        }
    } else if (cl < 0.0f && !b.smooth && t > 0.0f && m < 1.0f) {
        const float eta = b.eta;
        V3 h = vnorm(vneg(vadd(vscale(v, eta), l)));
        if (vdot(h, n) < 0.0f) h = vneg(h);
        const float vh = vdot(v, h), lh = vdot(l, h), nh = vdot(n, h);
        float den = eta * vh + lh;
        den *= den;
        if (vh > 0.0f && lh < 0.0f && nh > 0.0f && den > 1e-12f) {
            const float fd = fresnel_dielectric(vh, eta);
            const float d = ggx_d(nh, ggx_s2(n, h), a);
            const float g2 = ggx_vis(cv, cl, a) * 4.0f * cv * -cl;
            const float ft = vh * -lh * (1.0f - fd) * d * g2 / (cv * -cl * den);
            f = vscale(b.base, (1.0f - m) * t * ft);
            pdf = b.p[TRANS] * ggx_g1(cv, a) * vh * d / cv * -lh / den;
        }
    }
    return f;
}

__device__ __forceinline__ int bsdf_pick(const Bsdf &b, float ul)   // brdf.h:280-285
{
    int lobe = 0;
    for (float acc = b.p[0]; lobe < LOBES - 1 && (ul >= acc || b.p[lobe] <= 0.0f); acc += b.p[++lobe]);
    return b.p[lobe] > 0.0f ? lobe : -1;
}

struct BSample {
    V3 l, w;
    float pdf;
    bool delta;
};

__device__ bool bsdf_sample_lobe(Bsdf &b, const float *tab, int lobe, V3 n, V3 v, float u1, float u2, BSample &s)
{
    const float cv = vdot(v, n);
    if (cv <= 0.0f || lobe < 0) return false;
    bsdf_view(b, tab, cv);
    if (b.p[lobe] <= 0.0f) return false;
    const float m = b.metallic, t = b.transmission;
    s.delta = false;
    if (b.smooth && (lobe == SPEC || lobe == TRANS)) {
        const float fd = fresnel_dielectric(cv, b.eta);
        if (lobe == SPEC) {
            const V3 fm = fresnel_metal(b.base, cv);
            const float k = 1.0f / b.p[lobe];
            s.l = vreflect(vneg(v), n);
            s.w = v3(((1.0f - m) * fd + m * fm.x) * k, ((1.0f - m) * fd + m * fm.y) * k, ((1.0f - m) * fd + m * fm.z) * k);
        } else {
            if (fd >= 1.0f) return false;
            s.l = vrefract(vneg(v), n, b.eta);
            s.w = vscale(b.base, (1.0f - m) * t * (1.0f - fd) / b.p[lobe]);
        }
        s.pdf = 0.0f;
        s.delta = true;
        return true;
    }
    V3 l;
    if (lobe == DIFF || lobe == MS) l = cosine_hemisphere(n, u1, u2);
    else {
        const V3 h = ggx_sample_vndf(n, v, b.alpha, u1, u2);
        const float vh = vdot(v, h);
        if (vh <= 0.0f) return false;
        if (lobe == SPEC) l = vreflect(vneg(v), h);
        else {
            if (b.eta * b.eta * (1.0f - vh * vh) >= 1.0f) return false;
            l = vrefract(vneg(v), h, b.eta);
        }
    }
    if ((lobe == TRANS) != (vdot(l, n) < 0.0f)) return false;
    float pdf;
    const V3 f = bsdf_eval(b, tab, n, v, l, pdf);
    if (!(pdf > 0.0f)) return false;
    s.l = l;
    s.w = vscale(f, fabsf(vdot(l, n)) / pdf);
    s.pdf = pdf;
    return true;
}

__device__ __forceinline__ float power_heuristic(float pa, float pb)
{
    const float a = pa * pa, b = pb * pb;
    return a + b > 0.0f ? a / (a + b) : 0.0f;
}

// ---- appends to a queue: one atomic per wave ----
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *counter)
{
    const uint64_t mask = __ballot(want);
    if (mask == 0) return 0;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(mask & lanemask_lt());
}

// ---- kernels ----
// wf_gen (trace_wave.wgsl:51-70; trace_image tracer.c:335-346): each path's camera ray
__global__ __launch_bounds__(PBLOCK) void k_pt_gen(PtScene sc, PtJob job, PtBufs B, int P)
{
    const int p = blockIdx.x * PBLOCK + threadIdx.x;
    if (p == 0) B.count[0] = (uint32_t)P;
    if (p >= P) return;
    const int npix = job.W * job.H;
    const int pixel = p % npix;
    const uint32_t k = job.s0 + (uint32_t)(p / npix);
    const int x = pixel % job.W, y = pixel / job.W;
    const uint64_t seed = ((uint64_t)job.seed_hi << 32) | job.seed_lo;
    uint64_t st = pcg_seed((uint64_t)pixel, k, seed);
    const uint64_t inc = ((uint64_t)pixel << 1u) | 1u;
    const Smp sm{x, y, k, (uint32_t)seed * 0x9e3779b9u, sc.mask, &st, inc};
    float jx, jy;
    s2(sm, 0, jx, jy);
    const float u = ((float)x + jx) / (float)job.W, v = ((float)y + jy) / (float)job.H;
    // camera_ray (camera.h:38-42)
    const V3 org = v3(job.cam[0], job.cam[1], job.cam[2]);
    const V3 ul = v3(job.cam[3], job.cam[4], job.cam[5]), hz = v3(job.cam[6], job.cam[7], job.cam[8]),
             vt = v3(job.cam[9], job.cam[10], job.cam[11]);
    const V3 target = vsub(vadd(ul, vscale(hz, u)), vscale(vt, v));
    const V3 d = vnorm(vsub(target, org));
    B.qray[0][2 * (size_t)p] = make_float4(org.x, org.y, org.z, sc.eps);
    B.qray[0][2 * (size_t)p + 1] = make_float4(d.x, d.y, d.z, 1e30f);
    B.qpath[0][p] = (uint32_t)p;
    B.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    B.rad[p] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(EMIT_FULL));
    B.rng[p] = st;
}

// wf_ext (trace_wave.wgsl:72-84): closest hits of the ray queue `cur`; also clears the next
// queue's and the shadow queue's counters for the shade kernel that follows
__global__ __launch_bounds__(PBLOCK) void k_pt_extend(PtScene sc, PtBufs B, int cur, int P)
{
    __shared__ uint32_t stk[LR_LDS_STACK][PBLOCK];
    const int i = blockIdx.x * PBLOCK + threadIdx.x;
    if (i == 0) {
        B.count[1 - cur] = 0;
        B.count[2] = 0;
    }
    if (i >= (int)B.count[cur]) return;
    const float4 r0 = B.qray[cur][2 * (size_t)i], r1 = B.qray[cur][2 * (size_t)i + 1];
    const WalkStack S{&stk[0][threadIdx.x], PBLOCK, B.spill + i, (size_t)P, LR_LDS_STACK + B.spill_per};
    uint32_t cn, ct;
    const WalkHit h = walk<false, true, false>(sc.nodes, sc.root, sc.aux, sc.tris, sc.prims, sc.nprims, v3of(r0), v3of(r1),
                                               r0.w, r1.w, S, B.overflow, cn, ct);
    B.hits[i] = make_float4(h.t, __uint_as_float(h.id), h.b1, h.b2);
}

// wf_shade (trace_wave.wgsl:191-310), the bounce of trace_path_sampled (tracer.c:382-507)
__global__ __launch_bounds__(PBLOCK) void k_pt_shade(PtScene sc, PtJob job, PtBufs B, int cur, int depth)
{
    const int i = blockIdx.x * PBLOCK + threadIdx.x;
    const bool live = i < (int)B.count[cur];
    bool next = false, shadow = false;
    float4 nr0, nr1, sr0, sr1, scon;
    uint32_t p = 0;
    if (live) {
        p = B.qpath[cur][i];
        const float4 r0 = B.qray[cur][2 * (size_t)i], r1 = B.qray[cur][2 * (size_t)i + 1];
        const float4 hv = B.hits[i];
        float4 thr4 = B.thr[p], rad4 = B.rad[p];
        V3 thr = v3of(thr4), rad = v3of(rad4);
        float bounce_pdf = thr4.w;
        int emit = __float_as_int(rad4.w);
        const uint32_t id = __float_as_uint(hv.y);
        const V3 d = v3of(r1), o = v3of(r0);
        const int npix = job.W * job.H;
        const int pixel = (int)(p % (uint32_t)npix);
        const uint32_t k = job.s0 + p / (uint32_t)npix;
        uint64_t st = B.rng[p];
        const uint64_t inc = ((uint64_t)pixel << 1u) | 1u;
        const uint64_t seed = ((uint64_t)job.seed_hi << 32) | job.seed_lo;
        const Smp sm{pixel % job.W, pixel / job.W, k, (uint32_t)seed * 0x9e3779b9u, sc.mask, &st, inc};
        if (id != LR_NONE) {
            // the hit (set_hit, intersect.c:4-10): point, the normal facing the ray
            const float t = hv.x;
            const V3 hp = vadd(o, vscale(d, t));
            V3 outward;
            int mi;
            if (id < sc.nprims) {
                const Prim pr = load_prim(sc.prims, id);
                outward = pr.q.w < 0.5f ? vscale(vsub(hp, v3of(pr.c)), 1.0f / pr.c.w) : v3of(pr.n);
                mi = (int)pr.u.w;
            } else {
                const size_t ti = id - sc.nprims;
                const float4 tu = sc.tris[3 * ti + 1], tv = sc.tris[3 * ti + 2];
                outward = vnorm(vcross(v3of(tu), v3of(tv)));
                mi = (int)(__float_as_uint(tv.w) >> 26);
            }
            const bool front = vdot(d, outward) < 0.0f;
            const V3 n = front ? outward : vneg(outward);
            int type;
            Bsdf b = bsdf_make(sc.mats, mi, front, sc.tab, type);
            if (type == MAT_LIGHT) {
                if (front) {
                    const float4 em4 = sc.mats[3 * mi + 1];
                    const V3 em = v3of(em4);
                    if (emit == EMIT_FULL) rad = vadd(rad, vmul(thr, em));
                    else if (emit == EMIT_MIS && (int)id == sc.light) {
                        const Prim l = load_prim(sc.prims, (uint32_t)sc.light);
                        const float cos_light = -vdot(v3of(l.n), d);
                        if (cos_light > 0.0f) {
                            const float w = power_heuristic(bounce_pdf, (t * t) / (cos_light * l.v.w));
                            rad = vadd(rad, vmul(thr, vscale(em, w)));
                        }
                    }
                }
            } else if (depth < job.max_depth) {
                const V3 v = vneg(vnorm(d));
                // 1. next event estimation to the lamp (sample_direct_light, tracer.c:56-81),
                //    its visibility left to the shadow kernel
                const bool rough_part = !b.smooth || (1.0f - b.metallic) * (1.0f - b.transmission) > 0.0f;
                if (rough_part && sc.light >= 0) {
                    float la, lb;
                    s2(sm, 1 + 3 * depth, la, lb);
                    const float ua = lb, ub = la;   // the pair's second number first (tracer.c:50-53)
                    const Prim l = load_prim(sc.prims, (uint32_t)sc.light);
                    const V3 le = v3of(sc.mats[3 * (int)l.u.w + 1]);
                    const V3 xl = vadd(v3of(l.q), vadd(vscale(v3of(l.u), ua), vscale(v3of(l.v), ub)));
                    const V3 to_light = vsub(xl, hp);
                    const float dist2 = vdot(to_light, to_light);
                    const float dist = sqrtf(dist2);
                    const V3 wi = vscale(to_light, 1.0f / dist);
                    const float cos_surf = vdot(n, wi);
                    const float cos_light = -vdot(v3of(l.n), wi);
                    if (cos_light > 0.0f) {
                        float bpdf;
                        const V3 f = bsdf_eval(b, sc.tab, n, v, wi, bpdf);
                        if (vmax3(f) > 0.0f) {
                            float g = fabsf(cos_surf) * cos_light * l.v.w / dist2;
                            if (job.mis) g *= power_heuristic(dist2 / (cos_light * l.v.w), bpdf);
                            const V3 c = vmul(thr, vscale(vmul(f, le), g));
                            shadow = true;
                            sr0 = make_float4(hp.x, hp.y, hp.z, sc.eps);
                            sr1 = make_float4(wi.x, wi.y, wi.z, dist - sc.margin);
                            scon = make_float4(c.x, c.y, c.z, 0.0f);
                        }
                    }
                }
                // 2. the next direction from the BSDF, Russian roulette from depth 3
                bsdf_view(b, sc.tab, vdot(v, n));
                float ur = 0.0f, ul = 0.0f, da = 0.0f, db = 0.0f;
                const int nl = (b.p[0] > 0.0f) + (b.p[1] > 0.0f) + (b.p[2] > 0.0f) + (b.p[3] > 0.0f);
                if (depth >= 3 || nl > 1) s2(sm, 3 + 3 * depth, ur, ul);
                const int lobe = bsdf_pick(b, ul);
                if (lobe >= 0 && !(b.smooth && (lobe == SPEC || lobe == TRANS))) s2(sm, 2 + 3 * depth, da, db);
                BSample bs;
                if (bsdf_sample_lobe(b, sc.tab, lobe, n, v, da, db, bs)) {
                    thr = vmul(thr, bs.w);
                    if (bs.delta) emit = EMIT_FULL;
                    else {
                        bounce_pdf = bs.pdf;
                        emit = job.mis ? EMIT_MIS : EMIT_NONE;
                    }
                    bool go = true;
                    if (depth >= 3) {
                        const float pr = rfmin(vmax3(thr), 0.95f);
                        if (ur >= pr) go = false;
                        else thr = vscale(thr, 1.0f / pr);
                    }
                    if (go) {
                        const V3 nd = job.mis ? vnorm(bs.l) : bs.l;
                        next = true;
                        nr0 = make_float4(hp.x, hp.y, hp.z, sc.eps);
                        nr1 = make_float4(nd.x, nd.y, nd.z, 1e30f);
                    }
                }
            }
        }
        B.thr[p] = make_float4(thr.x, thr.y, thr.z, bounce_pdf);
        B.rad[p] = make_float4(rad.x, rad.y, rad.z, __int_as_float(emit));
        B.rng[p] = st;
    }
    const uint32_t slot = wave_append(next, &B.count[1 - cur]);
    if (next) {
        B.qray[1 - cur][2 * (size_t)slot] = nr0;
        B.qray[1 - cur][2 * (size_t)slot + 1] = nr1;
        B.qpath[1 - cur][slot] = p;
    }
    const uint32_t ss = wave_append(shadow, &B.count[2]);
    if (shadow) {
        B.sray[2 * (size_t)ss] = sr0;
        B.sray[2 * (size_t)ss + 1] = sr1;
        B.spath[ss] = p;
        B.scon[ss] = scon;
    }
}

// wf_shadow (trace_wave.wgsl:335-355): any hit before the lamp; if none, the contribution
__global__ __launch_bounds__(PBLOCK) void k_pt_shadow(PtScene sc, PtBufs B, int P)
{
    __shared__ uint32_t stk[LR_LDS_STACK][PBLOCK];
    const int i = blockIdx.x * PBLOCK + threadIdx.x;
    if (i >= (int)B.count[2]) return;
    const float4 r0 = B.sray[2 * (size_t)i], r1 = B.sray[2 * (size_t)i + 1];
    const WalkStack S{&stk[0][threadIdx.x], PBLOCK, B.spill + i, (size_t)P, LR_LDS_STACK + B.spill_per};
    uint32_t cn, ct;
    const WalkHit h = walk<true, true, false>(sc.nodes, sc.root, sc.aux, sc.tris, sc.prims, sc.nprims, v3of(r0), v3of(r1),
                                              r0.w, r1.w, S, B.overflow, cn, ct);
    if (h.id != LR_NONE) return;
    const uint32_t p = B.spath[i];
    const float4 c = B.scon[i];
    float4 r = B.rad[p];
    r.x += c.x;
    r.y += c.y;
    r.z += c.z;
    B.rad[p] = r;
}

// wf_resolve (trace_wave.wgsl:357-378): each pixel's samples of the pass, in sample order
// (NaNs skipped, the clamp applied: trace_image, tracer.c:347-349), into its (sum, count)
__global__ __launch_bounds__(PBLOCK) void k_pt_resolve(PtJob job, PtBufs B, float4 *accum)
{
    const int pixel = blockIdx.x * PBLOCK + threadIdx.x;
    const int npix = job.W * job.H;
    if (pixel >= npix) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int q = 0; q < job.S; q++) {
        const float4 r = B.rad[(size_t)q * npix + pixel];
        V3 c = v3(r.x, r.y, r.z);
        if (!(c.x == c.x && c.y == c.y && c.z == c.z)) continue;
        if (job.clamp > 0.0f && vmax3(c) > job.clamp) c = vscale(c, job.clamp / vmax3(c));
        sx += c.x;
        sy += c.y;
        sz += c.z;
    }
    float4 a = accum[pixel];
    a.x += sx;
    a.y += sy;
    a.z += sz;
    a.w += (float)job.S;
    accum[pixel] = a;
}

}  // namespace

// ---- host side -------------------------------------------------------------------------------
struct LrhPtScene {
    LrhContext *ctx;
    float4 *prims = nullptr, *tris = nullptr, *mats = nullptr, *boxes = nullptr;
    uint32_t *nodes = nullptr, *aux = nullptr;
    float *tab = nullptr;
    uint16_t *mask = nullptr;
    int nprims = 0, ntris = 0, nmats = 0, light = -1, depth = 1;
    float eps = 1e-4f, margin = 1e-3f;
};

static void pt_scene_release(LrhPtScene *s)
{
    if (!s) return;
    void *ps[] = {s->prims, s->tris, s->mats, s->boxes, s->nodes, s->aux, s->tab, s->mask};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete s;
}

extern "C" void lrs_blue_noise(uint16_t *rank);
extern "C" void lrs_brdf_tables(float *out);

static void quad_bounds(const float *p, float *lo, float *hi)
{
    // primitive_bounds (scene.c:47-63): the 4 corners (or the sphere's box), padded by 1e-3
    if (p[4 + 3] < 0.5f) {
        for (int k = 0; k < 3; k++) { lo[k] = p[k] - p[3]; hi[k] = p[k] + p[3]; }
    } else {
        const float *q = p + 4, *u = p + 8, *v = p + 12;
        for (int k = 0; k < 3; k++) {
            const float c[4] = {q[k], q[k] + u[k], q[k] + v[k], q[k] + (u[k] + v[k])};
            lo[k] = hi[k] = c[0];
            for (int j = 1; j < 4; j++) {
                lo[k] = c[j] != c[j] ? lo[k] : (lo[k] < c[j] ? lo[k] : c[j]);
                hi[k] = c[j] != c[j] ? hi[k] : (hi[k] > c[j] ? hi[k] : c[j]);
            }
        }
    }
    for (int k = 0; k < 3; k++) { lo[k] -= 1e-3f; hi[k] += 1e-3f; }
}

extern "C" int lrh_pt_scene(LrhContext *c, const float *prims, int nprims, const float *tris, int ntris,
                            const LrhMaterial *mats, int nmats, int light, float eps, float margin, int builder,
                            LrhPtScene **out)
{
    if (!c || !out || nprims < 0 || ntris < 0 || nprims + ntris <= 0 || nmats <= 0 || nmats > 64 || !mats ||
        (nprims > 0 && !prims) || (ntris > 0 && !tris) || light >= nprims)
        return lr_fail(LRH_EINVAL, "lrh_pt_scene: bad arguments");
    if (light >= 0 && prims[24 * light + 7] < 0.5f) return lr_fail(LRH_EINVAL, "lrh_pt_scene: the lamp must be a quad");
    for (int i = 0; i < nprims; i++) {
        const float m = prims[24 * (size_t)i + 11];
        if (!(m >= 0.0f && m < (float)nmats)) return lr_fail(LRH_EINVAL, "lrh_pt_scene: prim %d: material %g of %d", i, m, nmats);
    }
    for (int i = 0; i < ntris; i++) {
        uint32_t w;
        memcpy(&w, tris + 12 * (size_t)i + 11, 4);
        if ((int)(w >> 26) >= nmats) return lr_fail(LRH_EINVAL, "lrh_pt_scene: triangle %d: material %u of %d", i, w >> 26, nmats);
    }
    for (int i = 0; i < nmats; i++)
        if (mats[i].type < 0 || mats[i].type > LRH_MAT_PBR) return lr_fail(LRH_EINVAL, "lrh_pt_scene: material %d: type %d", i, mats[i].type);
    *out = nullptr;
    LrhPtScene *s = new LrhPtScene();
    s->ctx = c;
    s->nprims = nprims;
    s->ntris = ntris;
    s->nmats = nmats;
    s->light = light;
    s->eps = eps;
    s->margin = margin;
    const int n = nprims + ntris;
    std::vector<float> hm(12 * (size_t)nmats), boxes(8 * (size_t)n, 0.0f);
    for (int i = 0; i < nmats; i++) {
        const LrhMaterial &m = mats[i];
        float *o = &hm[12 * (size_t)i];
        o[0] = m.albedo[0]; o[1] = m.albedo[1]; o[2] = m.albedo[2]; o[3] = (float)m.type;
        o[4] = m.emission[0]; o[5] = m.emission[1]; o[6] = m.emission[2]; o[7] = m.ior;
        o[8] = m.metallic; o[9] = m.roughness; o[10] = m.transmission; o[11] = 0.0f;
    }
    for (int i = 0; i < nprims; i++) quad_bounds(prims + 24 * (size_t)i, &boxes[8 * (size_t)i], &boxes[8 * (size_t)i + 4]);
    for (int i = 0; i < ntris; i++) {
        const float *t = tris + 12 * (size_t)i;
        float *o = &boxes[8 * (size_t)(nprims + i)];
        for (int k = 0; k < 3; k++) {   // triangle_bounds (scene.c:65-74)
            const float q = t[k], a = q + t[4 + k], b = q + t[8 + k];
            float lo = q, hi = q;
            lo = a != a ? lo : (lo < a ? lo : a); hi = a != a ? hi : (hi > a ? hi : a);
            lo = b != b ? lo : (lo < b ? lo : b); hi = b != b ? hi : (hi > b ? hi : b);
            o[k] = lo - 1e-3f;
            o[4 + k] = hi + 1e-3f;
        }
    }
    std::vector<float> tab(2112);
    std::vector<uint16_t> mask(4096);
    lrs_brdf_tables(tab.data());
    lrs_blue_noise(mask.data());
    auto up = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0) bytes = 16;
        if (hipMalloc(dst, bytes) != hipSuccess) return false;
        if (src && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return false;
        return true;
    };
    bool ok = up((void **)&s->prims, nprims ? prims : nullptr, (size_t)nprims * 96) &&
              up((void **)&s->tris, ntris ? tris : nullptr, (size_t)ntris * 48) &&
              up((void **)&s->mats, hm.data(), hm.size() * 4) && up((void **)&s->boxes, boxes.data(), boxes.size() * 4) &&
              up((void **)&s->nodes, nullptr, (2 * (size_t)n - 1) * 32) && up((void **)&s->aux, nullptr, (size_t)n * 4) &&
              up((void **)&s->tab, tab.data(), tab.size() * 4) && up((void **)&s->mask, mask.data(), mask.size() * 2);
    if (!ok) {
        (void)hipGetLastError();
        pt_scene_release(s);
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene: device memory for the scene could not be had");
    }
    LrhBuildStats st;
    int rc = lrh_build(c, (const float *)s->boxes, n, nullptr, builder, 0, 0, 0, s->nodes, s->aux, nullptr, &st);
    if (rc) {
        pt_scene_release(s);
        return rc;
    }
    s->depth = st.depth;
    *out = s;
    return 0;
}

extern "C" void lrh_pt_scene_free(LrhPtScene *s)
{
    if (s && s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    pt_scene_release(s);
}

// the path buffers of a pass of P paths, in the context's path-tracing scratch
static int pt_bufs(LrhContext *c, int P, int spill_per, PtBufs &B)
{
    const size_t Pz = (size_t)P;
    size_t sizes[] = {Pz * 16, Pz * 16, Pz * 8, Pz * 32, Pz * 32, Pz * 4, Pz * 4, Pz * 16, Pz * 32, Pz * 4, Pz * 16, 256,
                      Pz * 4 * (size_t)spill_per};
    size_t total = 0;
    for (size_t z : sizes) total += lr_align(z);
    int rc = lr_scratch_reserve(c, LR_SCRATCH_PT, total);
    if (rc) return rc;
    char *at = (char *)c->scratch[LR_SCRATCH_PT].base;
    void *ptr[13];
    for (int i = 0; i < 13; i++) {
        ptr[i] = at;
        at += lr_align(sizes[i]);
    }
    B.thr = (float4 *)ptr[0];
    B.rad = (float4 *)ptr[1];
    B.rng = (uint64_t *)ptr[2];
    B.qray[0] = (float4 *)ptr[3];
    B.qray[1] = (float4 *)ptr[4];
    B.qpath[0] = (uint32_t *)ptr[5];
    B.qpath[1] = (uint32_t *)ptr[6];
    B.hits = (float4 *)ptr[7];
    B.sray = (float4 *)ptr[8];
    B.spath = (uint32_t *)ptr[9];
    B.scon = (float4 *)ptr[10];
    B.count = (uint32_t *)ptr[11];
    B.spill = (uint32_t *)ptr[12];
    B.spill_per = spill_per;
    B.overflow = c->flags;
    return 0;
}

int lr_spill_per(int max_depth);

extern "C" int lrh_pt_render(LrhContext *c, const LrhPtScene *s, const float *cam, int W, int H, int spp, uint32_t s0,
                             uint64_t seed, int max_depth, int mis, float clamp, float *d_accum)
{
    if (!c || !s || !cam || W <= 0 || H <= 0 || spp <= 0 || max_depth < 0 || !d_accum)
        return lr_fail(LRH_EINVAL, "lrh_pt_render: bad arguments");
    const int npix = W * H;
    // samples per pass: paths up to 2^22 at a time (about 600 MB of path state and queues)
    int S = (1 << 22) / npix;
    if (S < 1) S = 1;
    if (S > spp) S = spp;
    if ((int64_t)npix * S > (1 << 26)) return lr_fail(LRH_EREFUSED, "lrh_pt_render: %d x %d is too large", W, H);
    const int spill_per = lr_spill_per(s->depth);
    PtBufs B;
    int rc = pt_bufs(c, npix * S, spill_per, B);
    if (rc) return rc;
    PtScene sc{(const float4 *)s->nodes, s->aux, s->tris, s->prims, s->mats, s->tab, s->mask, 0, (uint32_t)s->nprims,
               s->light, s->eps, s->margin};
    PtJob job;
    for (int k = 0; k < 12; k++) job.cam[k] = cam[k];
    job.W = W;
    job.H = H;
    job.seed_lo = (uint32_t)seed;
    job.seed_hi = (uint32_t)(seed >> 32);
    job.max_depth = max_depth;
    job.mis = mis;
    job.clamp = clamp;
    for (int done = 0; done < spp; done += S) {
        job.S = spp - done < S ? spp - done : S;
        job.s0 = s0 + (uint32_t)done;
        const int P = npix * job.S;
        const dim3 grid((P + PBLOCK - 1) / PBLOCK), blk(PBLOCK);
        hipLaunchKernelGGL(k_pt_gen, grid, blk, 0, c->stream, sc, job, B, P);
        int cur = 0;
        for (int depth = 0; depth <= max_depth; depth++) {
            hipLaunchKernelGGL(k_pt_extend, grid, blk, 0, c->stream, sc, B, cur, P);
            hipLaunchKernelGGL(k_pt_shade, grid, blk, 0, c->stream, sc, job, B, cur, depth);
            hipLaunchKernelGGL(k_pt_shadow, grid, blk, 0, c->stream, sc, B, P);
            cur = 1 - cur;
        }
        hipLaunchKernelGGL(k_pt_resolve, dim3((npix + PBLOCK - 1) / PBLOCK), blk, 0, c->stream, job, B,
                           reinterpret_cast<float4 *>(d_accum));
        LR_HIP(hipGetLastError());
    }
    return 0;
}
