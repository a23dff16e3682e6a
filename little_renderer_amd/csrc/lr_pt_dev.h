// lr_pt_dev.h - the device side of the path tracer's light transport, shared by the engine's
// kernels: PCG32 (engine/base/rng.h), the Owen-scrambled Sobol sampler with its blue-noise
// shift (engine/base/sampler.c, trace.wgsl:81-126), and the BSDF (engine/scene/brdf.h:43-348,
// with the tables of brdf_tables.h), in the f32 arithmetic of the reference's float build.
#pragma once

#include "lr_walk.h"

namespace {

constexpr float PI_F = 3.14159265358979323846f;
constexpr int PAIRS = 13;                  // SAMPLER_PAIRS (sampler.h:29)
enum { EMIT_FULL = 0, EMIT_NONE = 1, EMIT_MIS = 2 };
enum { MAT_DIFFUSE = 0, MAT_MIRROR = 1, MAT_GLASS = 2, MAT_LIGHT = 3, MAT_PBR = 4 };
enum { SPEC = 0, DIFF = 1, TRANS = 2, MS = 3, LOBES = 4 };

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

// p[lobe] without dynamic indexing (the array then stays in registers)
__device__ __forceinline__ float lobe_p(const Bsdf &b, int lobe)
{
    return lobe == 0 ? b.p[0] : lobe == 1 ? b.p[1] : lobe == 2 ? b.p[2] : b.p[3];
}

__device__ __forceinline__ int bsdf_pick(const Bsdf &b, float ul)   // brdf.h:280-285
{
    int lobe = 0;
    // unrolled, so that the probabilities stay in registers: the sums in the same order
    float acc = b.p[0];
    if (ul >= acc || b.p[0] <= 0.0f) {
        lobe = 1;
        acc += b.p[1];
        if (ul >= acc || b.p[1] <= 0.0f) {
            lobe = 2;
            acc += b.p[2];
            if (ul >= acc || b.p[2] <= 0.0f) {
                lobe = 3;
                acc += b.p[3];
            }
        }
    }
    return lobe_p(b, lobe) > 0.0f ? lobe : -1;
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
    const float plobe = lobe_p(b, lobe);
    if (plobe <= 0.0f) return false;
    const float m = b.metallic, t = b.transmission;
    s.delta = false;
    if (b.smooth && (lobe == SPEC || lobe == TRANS)) {
        const float fd = fresnel_dielectric(cv, b.eta);
        if (lobe == SPEC) {
            const V3 fm = fresnel_metal(b.base, cv);
            const float k = 1.0f / plobe;
            s.l = vreflect(vneg(v), n);
            s.w = v3(((1.0f - m) * fd + m * fm.x) * k, ((1.0f - m) * fd + m * fm.y) * k, ((1.0f - m) * fd + m * fm.z) * k);
        } else {
            if (fd >= 1.0f) return false;
            s.l = vrefract(vneg(v), n, b.eta);
            s.w = vscale(b.base, (1.0f - m) * t * (1.0f - fd) / plobe);
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

}  // namespace

// ---- the sky (engine/scene/env.c:36-203; trace.wgsl:700-783) ----
namespace {
struct EnvRef {
    const float *rgb, *marg, *cond, *prob;   // w*h*3; h+1; h*(w+1); w*h (each pixel's probability)
    int w, h;
    float scale, rot, rot_c, rot_s;
};
constexpr float SKY_PI = 3.14159265358979323846f;

// env_uv + map_uv (env.c:36-57): map coordinates of direction d
__device__ __forceinline__ void env_uv(const EnvRef &e, V3 d, float &u, float &v)
{
    if (e.rot != 0.0f) d = v3(d.x * e.rot_c - d.z * e.rot_s, d.y, d.x * e.rot_s + d.z * e.rot_c);
    const float y = d.y < -1.0f ? -1.0f : d.y > 1.0f ? 1.0f : d.y;
    const float uu = 0.5f + atan2f(d.x, d.z) * (0.5f / SKY_PI);
    u = uu - floorf(uu);
    v = acosf(y) * (1.0f / SKY_PI);
}
// env_dir + map_dir (env.c:48-68)
__device__ __forceinline__ V3 env_dir(const EnvRef &e, float u, float v)
{
    const float ph = (u - 0.5f) * (2.0f * SKY_PI), th = v * SKY_PI;
    const float st = sinf(th);
    const V3 d = v3(st * sinf(ph), cosf(th), st * cosf(ph));
    return e.rot != 0.0f ? v3(d.x * e.rot_c + d.z * e.rot_s, d.y, -d.x * e.rot_s + d.z * e.rot_c) : d;
}
// bilerp (env.c:70-88): across wraps around, down stops at the poles
__device__ __forceinline__ V3 env_bilerp(const EnvRef &e, float u, float v)
{
    const int w = e.w, h = e.h;
    const float x = u * w - 0.5f, y = v * h - 0.5f;
    const int x0 = (int)floorf(x);
    int y0 = (int)floorf(y);
    const float fx = x - x0, fy = y - y0;
    int y1 = y0 + 1;
    if (y0 < 0) y0 = 0;
    if (y1 > h - 1) y1 = h - 1;
    if (y0 > h - 1) y0 = h - 1;
    int xa = x0 < 0 ? x0 + w : (x0 >= w ? x0 - w : x0);
    if (xa < 0 || xa >= w) xa = ((xa % w) + w) % w;
    const int xb = xa + 1 == w ? 0 : xa + 1;
    const float *a = e.rgb + 3 * ((size_t)y0 * w + xa), *b = e.rgb + 3 * ((size_t)y0 * w + xb);
    const float *c = e.rgb + 3 * ((size_t)y1 * w + xa), *dd = e.rgb + 3 * ((size_t)y1 * w + xb);
    return v3((1 - fy) * ((1 - fx) * a[0] + fx * b[0]) + fy * ((1 - fx) * c[0] + fx * dd[0]),
              (1 - fy) * ((1 - fx) * a[1] + fx * b[1]) + fy * ((1 - fx) * c[1] + fx * dd[1]),
              (1 - fy) * ((1 - fx) * a[2] + fx * b[2]) + fy * ((1 - fx) * c[2] + fx * dd[2]));
}
__device__ __forceinline__ V3 env_eval(const EnvRef &e, V3 d)
{
    float u, v;
    env_uv(e, d, u, v);
    return vscale(env_bilerp(e, u, v), e.scale);
}
// the last i in 0..n-1 with c[i] <= x (env.c:105-112)
__device__ __forceinline__ int env_upper_bound(const float *c, int n, float x)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (c[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}
// env_sample (env.c:171-189): a direction by the map's luminance x sin(theta), its density
__device__ __forceinline__ V3 env_sample(const EnvRef &e, float u1, float u2, V3 &d, float &pdf)
{
    pdf = 0.0f;
    const int y = env_upper_bound(e.marg, e.h, u2);
    const float m0 = e.marg[y], m1 = e.marg[y + 1];
    float dv = m1 > m0 ? (u2 - m0) / (m1 - m0) : 0.5f;
    const float *c = e.cond + (size_t)y * (e.w + 1);
    const int x = env_upper_bound(c, e.w, u1);
    const float c0 = c[x], c1 = c[x + 1];
    float du = c1 > c0 ? (u1 - c0) / (c1 - c0) : 0.5f;
    if (dv > 0.9999f) dv = 0.9999f;
    if (du > 0.9999f) du = 0.9999f;
    const float u = (x + du) / e.w, v = (y + dv) / e.h;
    const float st = sinf(v * SKY_PI);
    if (st < 1e-6f) return v3(0.0f, 0.0f, 0.0f);
    d = env_dir(e, u, v);
    pdf = e.prob[(size_t)y * e.w + x] * e.w * e.h / ((2.0f * SKY_PI * SKY_PI) * st);
    return env_eval(e, d);
}
__device__ __forceinline__ float env_pdf(const EnvRef &e, V3 d)   // env.c:191-203
{
    float u, v;
    env_uv(e, d, u, v);
    const float st = sinf(v * SKY_PI);
    if (st < 1e-6f) return 0.0f;
    int x = (int)(u * e.w), y = (int)(v * e.h);
    if (x > e.w - 1) x = e.w - 1;
    if (y > e.h - 1) y = e.h - 1;
    return e.prob[(size_t)y * e.w + x] * e.w * e.h / ((2.0f * SKY_PI * SKY_PI) * st);
}
}  // namespace
