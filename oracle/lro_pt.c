/* lro_pt.c - the oracle's path tracer: the sampler (engine/base/sampler.c), the BSDF
 * (engine/scene/brdf.h), next event estimation with MIS and the bounce loop
 * (engine/rt/tracer.c:30-81, 295-509) and the image loop (tracer.c:329-356), restated in
 * plain C with `real` = float, the reference's default build (-fsingle-precision-constant:
 * every literal below is a float). TEST INFRASTRUCTURE ONLY (see lro.h).
 *
 * One deliberate difference from trace_image: each (pixel, sample) has its own PCG32 stream
 * (the pixel's seed of tracer.c:339 xor the sample index times a constant), so that samples can
 * be traced in any order, on any device, alone or batched, and give the same numbers. The
 * device path tracer uses the same rule.
 *
 * Built with -DLRO_DOUBLE (liblro_d.so), `real` is double: the image oracle in FP64 at the
 * identical samples (the sampler's numbers and the PCG32 streams are integers, the same in
 * either build). */
#include "lro_pt.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef LRO_DOUBLE
/* the same code in double (oracle/liblro_d.so): the same random numbers, every operation of the
   light transport rounded to FP64 instead of FP32 (float literals are exact in double) */
typedef double real;
#define sqrtf sqrt
#define fabsf fabs
#define sinf sin
#define cosf cos
#define atan2f atan2
#define acosf acos
#define floorf floor
#define copysignf copysign
#else
typedef float real;
#endif
typedef struct { real x, y, z; } V;
static inline V v3(real x, real y, real z) { V r = {x, y, z}; return r; }
static inline V vadd(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V vsub(V a, V b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V vmul(V a, V b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V vscale(V a, real s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V vneg(V a) { return v3(-a.x, -a.y, -a.z); }
static inline real vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V vcross(V a, V b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline real rfmin(real a, real b) { return b != b ? a : (a < b ? a : b); }
static inline real rfmax(real a, real b) { return b != b ? a : (a > b ? a : b); }
static inline real vlen(V a) { return sqrtf(vdot(a, a)); }
static inline V vnorm(V a) { real l = vlen(a); return l > 0 ? vscale(a, 1.0f / l) : a; }
static inline real vmax(V a) { return rfmax(a.x, rfmax(a.y, a.z)); }
static inline V vreflect(V d, V n) { return vsub(d, vscale(n, 2.0f * vdot(d, n))); }
static inline V vrefract(V d, V n, real eta)
{
    real cos_i = rfmin(vdot(vneg(d), n), 1.0f);
    V perp = vscale(vadd(d, vscale(n, cos_i)), eta);
    V para = vscale(n, -sqrtf(fabsf(1.0f - vdot(perp, perp))));
    return vadd(perp, para);
}
static inline V vload(const float *p) { return v3(p[0], p[1], p[2]); }

#define PI_ 3.14159265358979323846

/* ---- the sampler (sampler.c:7-175) ---- */
uint32_t lro_sampler_hash(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
static inline uint32_t reverse_bits(uint32_t x)
{
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
static inline uint32_t owen_reversed(uint32_t x, uint32_t seed)   /* sampler.c:30-37 (Burley 2020) */
{
    x ^= x * 0x3d20adeau;
    x += seed;
    x *= (seed >> 16) | 1u;
    x ^= x * 0x05526c56u;
    x ^= x * 0x53a22864u;
    return reverse_bits(x);
}
uint32_t lro_sampler_owen(uint32_t x, uint32_t seed) { return owen_reversed(reverse_bits(x), seed); }
static inline uint32_t sobol1_reversed(uint32_t i)                 /* sampler.c:46-54 */
{
    i ^= (i >> 1) & 0x55555555u;
    i ^= (i >> 2) & 0x33333333u;
    i ^= (i >> 4) & 0x0f0f0f0fu;
    i ^= (i >> 8) & 0x00ff00ffu;
    i ^= i >> 16;
    return i;
}
uint32_t lro_sampler_sobol1(uint32_t i) { return reverse_bits(sobol1_reversed(i)); }

/* the blue-noise mask: void-and-cluster (sampler.c:86-132) */
#define BN 64
#define BN_N (BN * BN)
static void bn_update(float *energy, const float *kern, int p, float sign)
{
    int px = p % BN, py = p / BN;
    for (int y = 0; y < BN; y++) {
        const float *k = kern + ((y - py) & (BN - 1)) * BN;
        float *e = energy + y * BN;
        for (int x = 0; x < px; x++) e[x] += sign * k[x - px + BN];
        for (int x = px; x < BN; x++) e[x] += sign * k[x - px];
    }
}
static int bn_extreme(const float *energy, const unsigned char *on, int want_on, int want_max)
{
    int best = -1;
    for (int i = 0; i < BN_N; i++) {
        if (on[i] != want_on) continue;
        if (best < 0 || (want_max ? energy[i] > energy[best] : energy[i] < energy[best])) best = i;
    }
    return best;
}
void lro_blue_noise(uint16_t *rank)
{
    float *kern = malloc(BN_N * 4), *energy = malloc(BN_N * 4), *e0 = malloc(BN_N * 4);
    unsigned char on[BN_N], on0[BN_N];
    const double sigma = (double)1.9f;   /* the float build's literal (-fsingle-precision-constant) */
    for (int y = 0; y < BN; y++)
        for (int x = 0; x < BN; x++) {
            int dx = x < BN / 2 ? x : x - BN, dy = y < BN / 2 ? y : y - BN;
            kern[y * BN + x] = (float)exp(-(dx * dx + dy * dy) / (2 * sigma * sigma));
        }
    memset(on, 0, sizeof(on));
    memset(energy, 0, BN_N * 4);
    LroRng rng;
    lro_rng_seed(&rng, 0x5eedb10eULL, 7);
    const int ones = BN_N / 10;
    for (int n = 0; n < ones;) {
        int p = (int)(lro_rng_u32(&rng) % BN_N);
        if (on[p]) continue;
        on[p] = 1;
        bn_update(energy, kern, p, 1);
        n++;
    }
    for (int it = 0; it < BN_N; it++) {
        int c = bn_extreme(energy, on, 1, 1);
        on[c] = 0;
        bn_update(energy, kern, c, -1);
        int v = bn_extreme(energy, on, 0, 0);
        on[v] = 1;
        bn_update(energy, kern, v, 1);
        if (v == c) break;
    }
    memcpy(on0, on, sizeof(on));
    memcpy(e0, energy, BN_N * 4);
    for (int r = ones - 1; r >= 0; r--) {
        int c = bn_extreme(energy, on, 1, 1);
        on[c] = 0;
        bn_update(energy, kern, c, -1);
        rank[c] = (uint16_t)r;
    }
    memcpy(on, on0, sizeof(on));
    memcpy(energy, e0, BN_N * 4);
    for (int r = ones; r < BN_N; r++) {
        int v = bn_extreme(energy, on, 0, 0);
        on[v] = 1;
        bn_update(energy, kern, v, 1);
        rank[v] = (uint16_t)r;
    }
    free(kern); free(energy); free(e0);
}

typedef struct {
    LroRng *rng;
    uint32_t index, seed;
    int x, y, on;
    const uint16_t *mask;
} Smp;

static inline real bn_value(const uint16_t *m, int x, int y)
{
    return ((real)m[(y & (BN - 1)) * BN + (x & (BN - 1))] + 0.5f) * (real)(1.0 / BN_N);
}
static inline real unit24(uint32_t v) { return (real)(v >> 8) * (1.0f / 16777216.0f); }
static inline real wrap01(real t) { return t >= 1 ? t - 1 : t; }

/* sampler_2d (sampler.c:143-160) */
static void sampler_2d(Smp *s, int dim, real *a, real *b)
{
    if (!s->on || dim >= LRO_SAMPLER_PAIRS) {
        *a = lro_rng_float(s->rng);
        *b = lro_rng_float(s->rng);
        return;
    }
    uint32_t tile = lro_sampler_hash(((uint32_t)(s->x >> 6) * 73856093u) ^ ((uint32_t)(s->y >> 6) * 19349663u));
    uint32_t seed = lro_sampler_hash(s->seed ^ lro_sampler_hash(((uint32_t)dim * 0x9e3779b9u) ^ tile));
    uint32_t i = lro_sampler_owen(s->index, seed);
    uint32_t xa = owen_reversed(i, lro_sampler_hash(seed ^ 0xa511e9b3u));
    uint32_t xb = owen_reversed(sobol1_reversed(i), lro_sampler_hash(seed ^ 0x63d83595u));
    int ox = s->x + 23 * dim, oy = s->y + 41 * dim;
    *a = wrap01(unit24(xa) + bn_value(s->mask, ox, oy));
    *b = wrap01(unit24(xb) + bn_value(s->mask, ox + 32, oy + 17));
}

void lro_sampler_2d(const uint16_t *mask, int x, int y, uint32_t index, uint32_t seed, int dim, float out[2])
{
    LroRng rng;
    lro_rng_seed(&rng, 1, 1);
    Smp s = {&rng, index, seed, x, y, 1, mask};
    real a, b;
    sampler_2d(&s, dim, &a, &b);
    out[0] = (float)a;
    out[1] = (float)b;
}

/* ---- the BSDF (brdf.h:43-348) ---- */
enum { SPEC, DIFF, TRANS, MS, LOBES };
typedef struct {
    V base;
    real metallic, transmission, rough, alpha, eta, e_avg, es_avg;
    int smooth;
    V ms;
    real kd_norm, kms_norm;
    real cv, p[LOBES], es_v, e_v;
    const float *tab;      /* E[32][32], B[32][32], E_avg[32], B_avg[32] */
} Bsdf;
typedef struct { V l, w; real pdf; int delta, through; } BSample;

static inline real clamp01(real x) { return x < 0 ? 0 : x > 1 ? 1 : x; }
static inline real lum(V c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
static inline real table(const float *t, real rough, real mu)      /* brdf.h:71-80 */
{
    real x = rough * (LRO_BRDF_N - 1), y = clamp01(mu) * (LRO_BRDF_N - 1);
    int j = (int)x, i = (int)y;
    if (j > LRO_BRDF_N - 2) j = LRO_BRDF_N - 2;
    if (i > LRO_BRDF_N - 2) i = LRO_BRDF_N - 2;
    real fx = x - j, fy = y - i;
    real e0 = t[j * 32 + i] + fy * (t[j * 32 + i + 1] - t[j * 32 + i]);
    real e1 = t[(j + 1) * 32 + i] + fy * (t[(j + 1) * 32 + i + 1] - t[(j + 1) * 32 + i]);
    return e0 + fx * (e1 - e0);
}
static inline real albedo_e(const Bsdf *b, real mu) { return table(b->tab, b->rough, mu); }
static inline real albedo_b(const Bsdf *b, real mu) { return table(b->tab + 1024, b->rough, mu); }
static inline real fresnel_dielectric(real cos_i, real eta)        /* brdf.h:88-98 */
{
    cos_i = clamp01(cos_i);
    real s2 = eta * eta * (1 - cos_i * cos_i);
    if (s2 >= 1) return 1;
    real c = eta > 1 ? sqrtf(1 - s2) : cos_i;
    real f0 = (1 - eta) / (1 + eta);
    f0 *= f0;
    real x = 1 - c, x2 = x * x;
    return f0 + (1 - f0) * x2 * x2 * x;
}
static inline real fresnel_dielectric_avg(real eta)               /* brdf.h:103-109 */
{
    real e = eta > 1 ? 1 / eta : eta;
    real f0 = (1 - e) / (1 + e);
    f0 *= f0;
    real avg = f0 + (1 - f0) / 21;
    return eta > 1 ? 1 - (1 - avg) / (eta * eta) : avg;
}
static inline real spec_albedo(const Bsdf *b, real mu)            /* brdf.h:114-119 */
{
    if (b->smooth || b->eta > 1) return fresnel_dielectric(mu, b->eta);
    real f0 = (1 - b->eta) / (1 + b->eta), e = albedo_e(b, mu), bb = albedo_b(b, mu);
    f0 *= f0;
    return f0 * (e - bb) + bb;
}
static inline real spec_albedo_avg(const Bsdf *b)
{
    return b->smooth || b->eta > 1 ? fresnel_dielectric_avg(b->eta) : b->es_avg;
}
static inline V fresnel_metal(V f0, real c)                       /* brdf.h:125-128 */
{
    real x = 1 - clamp01(c), x2 = x * x, x5 = x2 * x2 * x;
    return v3(f0.x + (1 - f0.x) * x5, f0.y + (1 - f0.y) * x5, f0.z + (1 - f0.z) * x5);
}
static inline real ggx_d(real nh, real s2, real a)                /* brdf.h:132-135 */
{
    real a2 = a * a, d = s2 + a2 * nh * nh;
    return a2 / ((real)PI_ * d * d);
}
static inline real ggx_s2(V n, V h) { V c = vcross(n, h); return vdot(c, c); }
static inline real ggx_g1(real c, real a)
{
    c = fabsf(c);
    return 2 * c / (c + sqrtf(a * a + (1 - a * a) * c * c));
}
static inline real ggx_vis(real cv, real cl, real a)
{
    cv = fabsf(cv); cl = fabsf(cl);
    real a2 = a * a;
    real d = cl * sqrtf(cv * cv * (1 - a2) + a2) + cv * sqrtf(cl * cl * (1 - a2) + a2);
    return d > 0 ? 0.5f / d : 0;
}
static void onb(V n, V *t, V *b)                                  /* rng.h:53-59 */
{
    real s = copysignf(1.0f, n.z);
    real a = -1.0f / (s + n.z);
    real c = n.x * n.y * a;
    *t = v3(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    *b = v3(c, s + n.y * n.y * a, -n.y);
}
static V cosine_hemisphere(V n, real u1, real u2)                /* rng.h:63-70 */
{
    real phi = 2.0f * (real)PI_ * u1;
    real rad = sqrtf(u2);
    real x = rad * cosf(phi), y = rad * sinf(phi), z = sqrtf(rfmax(0.0f, 1.0f - u2));
    V t, b;
    onb(n, &t, &b);
    return vadd(vadd(vscale(t, x), vscale(b, y)), vscale(n, z));
}
static V ggx_sample_vndf(V n, V v, real a, real u1, real u2)     /* brdf.h:154-164 */
{
    V t, bt;
    onb(n, &t, &bt);
    V vh = vnorm(v3(a * vdot(v, t), a * vdot(v, bt), vdot(v, n)));
    real phi = 2 * (real)PI_ * u1;
    real z = (1 - u2) * (1 + vh.z) - vh.z;
    real st = sqrtf(rfmax(0, 1 - z * z));
    V hs = v3(st * cosf(phi) + vh.x, st * sinf(phi) + vh.y, z + vh.z);
    V hl = vnorm(v3(a * hs.x, a * hs.y, rfmax(hs.z, 0)));
    return vadd(vadd(vscale(t, hl.x), vscale(bt, hl.y)), vscale(n, hl.z));
}
static V ms_color(const Bsdf *b)                                  /* brdf.h:168-172 */
{
    real e = b->e_avg;
    V fa = v3(b->base.x + (1 - b->base.x) / 21, b->base.y + (1 - b->base.y) / 21, b->base.z + (1 - b->base.z) / 21);
    return v3(fa.x * fa.x * e / (1 - fa.x * (1 - e)), fa.y * fa.y * e / (1 - fa.y * (1 - e)), fa.z * fa.z * e / (1 - fa.z * (1 - e)));
}

/* material_resolved (scene.h:80-90) then bsdf_make (brdf.h:176-205) */
static Bsdf bsdf_make(const LroMat *mat, int outside, const float *tab)
{
    real metallic = mat->metallic, roughness = mat->roughness, ior = mat->ior, transmission = mat->transmission;
    switch (mat->type) {
    case LRO_MAT_DIFFUSE: metallic = 0; roughness = 1; ior = 1.5f; transmission = 0; break;
    case LRO_MAT_MIRROR: metallic = 1; roughness = 0; ior = 1.5f; transmission = 0; break;
    case LRO_MAT_GLASS: metallic = 0; roughness = 0; transmission = 1; if (!(ior > 0)) ior = 1.5f; break;
    case LRO_MAT_LIGHT: metallic = 0; roughness = 1; transmission = 0; break;
    default: if (!(ior > 0)) ior = 1.5f; break;
    }
    Bsdf b;
    b.tab = tab;
    b.base = vload(mat->albedo);
    b.metallic = clamp01(metallic);
    b.transmission = clamp01(transmission);
    real r = clamp01(roughness);
    b.smooth = r <= 0.02f;
    b.rough = r;
    b.alpha = rfmax(r * r, 1e-4f);
    ior = ior > 0 ? ior : 1.5f;
    b.eta = b.transmission > 0 && !outside ? ior : 1 / ior;
    real x = r * (LRO_BRDF_N - 1);
    int j = (int)x;
    if (j > LRO_BRDF_N - 2) j = LRO_BRDF_N - 2;
    const float *EA = tab + 2048, *BA = tab + 2080;
    b.e_avg = EA[j] + (x - j) * (EA[j + 1] - EA[j]);
    real b_avg = BA[j] + (x - j) * (BA[j + 1] - BA[j]);
    real f0 = (1 - b.eta) / (1 + b.eta);
    f0 *= f0;
    b.es_avg = f0 * (b.e_avg - b_avg) + b_avg;
    b.kd_norm = 1 / (1 - spec_albedo_avg(&b));
    b.kms_norm = 1 / ((real)PI_ * (1 - b.e_avg));
    b.ms = b.smooth || b.metallic <= 0 ? v3(0, 0, 0) : ms_color(&b);
    b.cv = -1;
    b.es_v = b.e_v = 0;
    for (int i = 0; i < LOBES; i++) b.p[i] = 0;
    return b;
}

static void bsdf_view(Bsdf *b, real cv)                           /* brdf.h:209-224 */
{
    if (b->cv == cv) return;
    b->cv = cv;
    real fd = fresnel_dielectric(cv, b->eta);
    real fs = b->smooth ? fd : rfmin(fd, 0.9f);
    real m = b->metallic, t = b->transmission, lb = lum(b->base);
    b->es_v = (1 - m) * (1 - t) > 0 ? spec_albedo(b, cv) : 0;
    b->e_v = b->smooth || m <= 0 ? 1 : albedo_e(b, cv);
    real w[LOBES];
    w[SPEC] = (1 - m) * fs + m * lum(fresnel_metal(b->base, cv));
    w[DIFF] = (1 - m) * (1 - t) * (1 - b->es_v) * lb;
    w[TRANS] = (1 - m) * t * (1 - fs) * lb;
    w[MS] = m * lum(b->ms) * (1 - b->e_v);
    real sum = w[0] + w[1] + w[2] + w[3];
    for (int k = 0; k < LOBES; k++) b->p[k] = sum > 0 ? w[k] / sum : 0;
}

static V bsdf_eval(Bsdf *b, V n, V v, V l, real *pdf)             /* brdf.h:228-275 */
{
    V f = v3(0, 0, 0);
    *pdf = 0;
    real cv = vdot(v, n), cl = vdot(l, n);
    if (cv <= 0) return f;
    bsdf_view(b, cv);
    const real *p = b->p;
    const real m = b->metallic, t = b->transmission, a = b->alpha;
    if (cl > 0) {
        V h = vnorm(vadd(v, l));
        real vh = clamp01(vdot(v, h)), nh = vdot(n, h);
        real fd = fresnel_dielectric(vh, b->eta);
        real kd = (1 - m) * (1 - t);
        if (kd > 0) {
            kd *= (1 - b->es_v) * (1 - spec_albedo(b, cl)) * b->kd_norm;
            f = vscale(b->base, kd / (real)PI_);
            *pdf += p[DIFF] * cl / (real)PI_;
        }
        if (!b->smooth && m > 0) {
            real k = m * (1 - b->e_v) * (1 - albedo_e(b, cl)) * b->kms_norm;
            f = vadd(f, vscale(b->ms, k));
            *pdf += p[MS] * cl / (real)PI_;
        }
        if (!b->smooth && nh > 0) {
            real d = ggx_d(nh, ggx_s2(n, h), a);
            V fm = fresnel_metal(b->base, vh);
            real s = d * ggx_vis(cv, cl, a);
            f = vadd(f, v3(((1 - m) * fd + m * fm.x) * s, ((1 - m) * fd + m * fm.y) * s, ((1 - m) * fd + m * fm.z) * s));
            *pdf += p[SPEC] * ggx_g1(cv, a) * d / (4 * cv);
        }
    } else if (cl < 0 && !b->smooth && t > 0 && m < 1) {
        const real eta = b->eta;
        V h = vnorm(vneg(vadd(vscale(v, eta), l)));
        if (vdot(h, n) < 0) h = vneg(h);
        real vh = vdot(v, h), lh = vdot(l, h), nh = vdot(n, h);
        real den = eta * vh + lh;
        den *= den;
        if (vh > 0 && lh < 0 && nh > 0 && den > 1e-12f) {
            real fd = fresnel_dielectric(vh, eta);
            real d = ggx_d(nh, ggx_s2(n, h), a);
            real g2 = ggx_vis(cv, cl, a) * 4 * cv * -cl;
            real ft = vh * -lh * (1 - fd) * d * g2 / (cv * -cl * den);
            f = vscale(b->base, (1 - m) * t * ft);
            *pdf = p[TRANS] * ggx_g1(cv, a) * vh * d / cv * -lh / den;
        }
    }
    return f;
}

static int bsdf_pick(const Bsdf *b, real ul)                      /* brdf.h:280-285 */
{
    const real *p = b->p;
    int lobe = 0;
    for (real acc = p[0]; lobe < LOBES - 1 && (ul >= acc || p[lobe] <= 0); acc += p[++lobe]);
    return p[lobe] > 0 ? lobe : -1;
}
static int bsdf_lobe_delta(const Bsdf *b, int lobe) { return b->smooth && (lobe == SPEC || lobe == TRANS); }
static int bsdf_nlobes(const Bsdf *b) { return (b->p[0] > 0) + (b->p[1] > 0) + (b->p[2] > 0) + (b->p[3] > 0); }

static int bsdf_sample_lobe(Bsdf *b, int lobe, V n, V v, real u1, real u2, BSample *s)   /* brdf.h:292-340 */
{
    real cv = vdot(v, n);
    if (cv <= 0 || lobe < 0) return 0;
    bsdf_view(b, cv);
    const real *p = b->p;
    if (p[lobe] <= 0) return 0;
    const real m = b->metallic, t = b->transmission;
    s->delta = 0;
    s->through = 0;
    if (b->smooth && (lobe == SPEC || lobe == TRANS)) {
        real fd = fresnel_dielectric(cv, b->eta);
        if (lobe == SPEC) {
            V fm = fresnel_metal(b->base, cv);
            real k = 1 / p[lobe];
            s->l = vreflect(vneg(v), n);
            s->w = v3(((1 - m) * fd + m * fm.x) * k, ((1 - m) * fd + m * fm.y) * k, ((1 - m) * fd + m * fm.z) * k);
        } else {
            if (fd >= 1) return 0;
            s->l = vrefract(vneg(v), n, b->eta);
            s->w = vscale(b->base, (1 - m) * t * (1 - fd) / p[lobe]);
            s->through = 1;
        }
        s->pdf = 0;
        s->delta = 1;
        return 1;
    }
    V l;
    if (lobe == DIFF || lobe == MS) l = cosine_hemisphere(n, u1, u2);
    else {
        V h = ggx_sample_vndf(n, v, b->alpha, u1, u2);
        real vh = vdot(v, h);
        if (vh <= 0) return 0;
        if (lobe == SPEC) l = vreflect(vneg(v), h);
        else {
            if (b->eta * b->eta * (1 - vh * vh) >= 1) return 0;
            l = vrefract(vneg(v), h, b->eta);
        }
    }
    if ((lobe == TRANS) != (vdot(l, n) < 0)) return 0;
    real pdf;
    V f = bsdf_eval(b, n, v, l, &pdf);
    if (!(pdf > 0)) return 0;
    s->l = l;
    s->w = vscale(f, fabsf(vdot(l, n)) / pdf);
    s->pdf = pdf;
    s->through = vdot(l, n) < 0;
    return 1;
}

void lro_bsdf_case(const LroMat *m, int outside, const float *tab, const float n[3], const float v[3], const float l[3],
                   float ul, float u1, float u2, float out[16])
{
    Bsdf b = bsdf_make(m, outside, tab);
    V N = vload(n), Vv = vload(v), L = vload(l);
    real pdf;
    V f = bsdf_eval(&b, N, Vv, L, &pdf);
    out[0] = f.x; out[1] = f.y; out[2] = f.z; out[3] = pdf;
    BSample s;
    memset(&s, 0, sizeof(s));
    bsdf_view(&b, vdot(Vv, N));
    int lobe = bsdf_pick(&b, ul);
    int ok = bsdf_sample_lobe(&b, lobe, N, Vv, u1, u2, &s);
    out[4] = (float)ok; out[5] = (float)lobe;
    out[6] = s.l.x; out[7] = s.l.y; out[8] = s.l.z;
    out[9] = s.w.x; out[10] = s.w.y; out[11] = s.w.z;
    out[12] = s.pdf; out[13] = (float)s.delta; out[14] = (float)s.through; out[15] = 0;
}

/* lro_bsdf_case over n cases of 17 floats (base rgb, metallic, roughness, ior, transmission,
 * outside, v xyz, l xyz, ul, u1, u2; material_pbr, scene.h:70-74; the normal +z): 16 floats each */
void lro_bsdf_cases(const float *in, int n, const float *tab, float *out)
{
    const float nz[3] = {0, 0, 1};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        const float *c = in + 17 * (size_t)i;
        LroMat m;
        memset(&m, 0, sizeof(m));
        m.type = LRO_MAT_PBR;
        m.albedo[0] = c[0]; m.albedo[1] = c[1]; m.albedo[2] = c[2];
        m.metallic = c[3]; m.roughness = c[4]; m.ior = c[5]; m.transmission = c[6];
        lro_bsdf_case(&m, c[7] > 0.5f, tab, nz, c + 8, c + 11, c[14], c[15], c[16], out + 16 * (size_t)i);
    }
}

/* ---- the scene: shapes and their hits (intersect.h:66-112, intersect.c:4-67) ---- */
typedef struct { real t; V p, n, ng; int front, prim, inst; } Hit;

static int hit_sphere(const LroPrim *s, V o, V d, real tmin, real tmax, real *t_out)
{
    V oc = vsub(vload(s->c), o);
    real a = vdot(d, d), hh = vdot(d, oc), c = vdot(oc, oc) - s->c[3] * s->c[3];
    real disc = hh * hh - a * c;
    if (disc < 0) return 0;
    real sq = sqrtf(disc);
    real t = (hh - sq) / a;
    if (t <= tmin || t >= tmax) {
        t = (hh + sq) / a;
        if (t <= tmin || t >= tmax) return 0;
    }
    *t_out = t;
    return 1;
}
static int hit_quad(const LroPrim *s, V o, V d, real tmin, real tmax, real *t_out)
{
    V nn = vload(s->n);
    real denom = vdot(nn, d);
    if (fabsf(denom) < 1e-12f) return 0;
    real t = (s->n[3] - vdot(nn, o)) / denom;
    if (t <= tmin || t >= tmax) return 0;
    V p = vadd(o, vscale(d, t));
    V rel = vsub(p, vload(s->q));
    real a = vdot(vload(s->w), vcross(rel, vload(s->v)));
    real b = vdot(vload(s->w), vcross(vload(s->u), rel));
    if (a < 0 || a > 1 || b < 0 || b > 1) return 0;
    *t_out = t;
    return 1;
}
static int hit_tri(const float *tr, V o, V d, real tmin, real tmax, real *t_out)
{
    V q = vload(tr), u = vload(tr + 4), v = vload(tr + 8);
    V pvec = vcross(d, v);
    real det = vdot(u, pvec);
    if (fabsf(det) < 1e-12f) return 0;
    real inv = 1.0f / det;
    V tvec = vsub(o, q);
    real b1 = vdot(tvec, pvec) * inv;
    if (b1 < 0 || b1 > 1) return 0;
    V qvec = vcross(tvec, u);
    real b2 = vdot(d, qvec) * inv;
    if (b2 < 0 || b1 + b2 > 1) return 0;
    real t = vdot(v, qvec) * inv;
    if (t <= tmin || t >= tmax) return 0;
    *t_out = t;
    return 1;
}
static int shape_test(const LroScene *s, int id, V o, V d, real tmin, real tmax, real *t)
{
    if (id < s->nprims) {
        const LroPrim *p = &s->prims[id];
        return p->q[3] < 0.5f ? hit_sphere(p, o, d, tmin, tmax, t) : hit_quad(p, o, d, tmin, tmax, t);
    }
    return hit_tri(s->tris + 12 * (size_t)(id - s->nprims), o, d, tmin, tmax, t);
}
/* ---- instances (trace.wgsl:186-262: to_object, hit_unit_sphere, hit_unit_box,
   normal_to_world) ---- */
static void to_object(const LroInst *in, V o, V d, V *oo, V *od)
{
    const float *a = in->inv, *b = in->inv + 4, *c = in->inv + 8;
    *oo = v3(a[0] * o.x + a[1] * o.y + a[2] * o.z + a[3], b[0] * o.x + b[1] * o.y + b[2] * o.z + b[3],
             c[0] * o.x + c[1] * o.y + c[2] * o.z + c[3]);
    *od = v3(a[0] * d.x + a[1] * d.y + a[2] * d.z, b[0] * d.x + b[1] * d.y + b[2] * d.z,
             c[0] * d.x + c[1] * d.y + c[2] * d.z);
}
/* an object-space normal in the world: by the inverse transpose, normalised */
static V normal_to_world(const LroInst *in, V n)
{
    const float *a = in->inv, *b = in->inv + 4, *c = in->inv + 8;
    return vnorm(v3(a[0] * n.x + b[0] * n.y + c[0] * n.z, a[1] * n.x + b[1] * n.y + c[1] * n.z,
                    a[2] * n.x + b[2] * n.y + c[2] * n.z));
}
static int hit_unit_sphere(V o, V d, real tmin, real tmax, real *t_out)
{
    real a = vdot(d, d), hh = -vdot(d, o), c = vdot(o, o) - 1.0f;
    real disc = hh * hh - a * c;
    if (disc < 0) return 0;
    real sq = sqrtf(disc);
    real t = (hh - sq) / a;
    if (t <= tmin || t >= tmax) {
        t = (hh + sq) / a;
        if (t <= tmin || t >= tmax) return 0;
    }
    *t_out = t;
    return 1;
}
static inline real safe_inv(real x) { return fabsf(x) > 1e-12f ? 1.0f / x : 1e30f; }
static int hit_unit_box(V o, V d, real tmin, real tmax, real *t_out)
{
    V inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    real t0x = (-1.0f - o.x) * inv.x, t0y = (-1.0f - o.y) * inv.y, t0z = (-1.0f - o.z) * inv.z;
    real t1x = (1.0f - o.x) * inv.x, t1y = (1.0f - o.y) * inv.y, t1z = (1.0f - o.z) * inv.z;
    real ax = t0x < t1x ? t0x : t1x, ay = t0y < t1y ? t0y : t1y, az = t0z < t1z ? t0z : t1z;
    real bx = t0x > t1x ? t0x : t1x, by = t0y > t1y ? t0y : t1y, bz = t0z > t1z ? t0z : t1z;
    real enter = ax > ay ? ax : ay;
    enter = enter > az ? enter : az;
    real leave = bx < by ? bx : by;
    leave = leave < bz ? leave : bz;
    if (leave < enter) return 0;
    if (enter > tmin && enter < tmax) { *t_out = enter; return 1; }
    if (leave > tmin && leave < tmax) { *t_out = leave; return 1; }
    return 0;
}
/* the object-space outward normal of the unit box at p: the axis p lies farthest along */
static V unit_box_normal(V p)
{
    real ax = fabsf(p.x), ay = fabsf(p.y), az = fabsf(p.z);
    if (ax >= ay && ax >= az) return v3(p.x < 0 ? -1.0f : 1.0f, 0, 0);
    if (ay >= az) return v3(0, p.y < 0 ? -1.0f : 1.0f, 0);
    return v3(0, 0, p.z < 0 ? -1.0f : 1.0f);
}
/* does the ray (object space) meet the padded box b in [tmin, tmax]? in double, so that the
   test cannot drop a hit the triangle test keeps */
static int box_meets(const float *b, V o, V d, real tmin, real tmax)
{
    double lo = tmin, hi = tmax;
    const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    for (int a = 0; a < 3; a++) {
        const double pad = 1e-3 * ((double)b[3 + a] - b[a]) + 1e-6;
        const double bl = b[a] - pad, bh = b[3 + a] + pad;
        if (fabs(dd[a]) < 1e-30) {
            if (oo[a] < bl || oo[a] > bh) return 0;
            continue;
        }
        double t0 = (bl - oo[a]) / dd[a], t1 = (bh - oo[a]) / dd[a];
        if (t0 > t1) { const double x = t0; t0 = t1; t1 = x; }
        if (t0 > lo) lo = t0;
        if (t1 < hi) hi = t1;
        if (lo > hi * (1 + 1e-9) + 1e-9) return 0;
    }
    return 1;
}

/* the closest (any: the first) hit among instance k's shapes, in its space: local id */
static int inst_test(const LroScene *s, int k, V o, V d, real tmin, real tmax, int any, real *t, int *local)
{
    const LroInst *in = &s->inst[k];
    V oo, od;
    to_object(in, o, d, &oo, &od);
    if (in->kind >= 0 && s->mesh_box && !box_meets(s->mesh_box + 6 * (size_t)in->kind, oo, od, tmin, tmax)) return 0;
    if (in->kind == LRO_INST_SPHERE || in->kind == LRO_INST_BOX) {
        real tt;
        int h = in->kind == LRO_INST_SPHERE ? hit_unit_sphere(oo, od, tmin, tmax, &tt) : hit_unit_box(oo, od, tmin, tmax, &tt);
        if (h) { *t = tt; *local = in->kind; }
        return h;
    }
    int found = 0;
    for (int i = s->mesh_first[in->kind]; i < s->mesh_first[in->kind + 1]; i++) {
        real tt;
        if (hit_tri(s->mtris + 12 * (size_t)i, oo, od, tmin, tmax, &tt)) {
            found = 1;
            tmax = tt;
            *t = tt;
            *local = i - s->mesh_first[in->kind];
            if (any) return 1;
        }
    }
    return found;
}

/* the slab test of a node's box against a ray, (tmin, tmax) */
static int node_meets(const uint32_t *nd, V o, V inv, real tmin, real tmax)
{
    const float *lo = (const float *)nd, *hi = (const float *)(nd + 4);
    real t0 = (lo[0] - o.x) * inv.x, t1 = (hi[0] - o.x) * inv.x;
    real a = t0 < t1 ? t0 : t1, b = t0 < t1 ? t1 : t0;
    t0 = (lo[1] - o.y) * inv.y; t1 = (hi[1] - o.y) * inv.y;
    a = rfmax(a, t0 < t1 ? t0 : t1); b = rfmin(b, t0 < t1 ? t1 : t0);
    t0 = (lo[2] - o.z) * inv.z; t1 = (hi[2] - o.z) * inv.z;
    a = rfmax(a, t0 < t1 ? t0 : t1); b = rfmin(b, t0 < t1 ? t1 : t0);
    return rfmax(a, tmin) <= rfmin(b, tmax);
}
/* the triangles through the tree: the closest (any: the first) in (tmin, *tmax), its index in
   *found; every box meeting the ray is opened (no order), so the kept hit is the brute force's
   but for equal t */
static int tree_tris(const LroScene *s, V o, V d, real tmin, real *tmax, int any, int *found)
{
    uint32_t stack[512];
    int sp = 0, hit = 0;
    const V inv = v3(fabsf(d.x) > 1e-30f ? 1 / d.x : 1e30f, fabsf(d.y) > 1e-30f ? 1 / d.y : 1e30f,
                     fabsf(d.z) > 1e-30f ? 1 / d.z : 1e30f);
    stack[sp++] = 0;
    while (sp > 0) {
        const uint32_t *nd = s->tnodes + 8 * (size_t)stack[--sp];
        if (!node_meets(nd, o, inv, tmin, *tmax)) continue;
        if (nd[7] & 0x80000000u) {
            for (uint32_t k = 0; k < (nd[7] & 0x7fffffffu); k++) {
                const int i = (int)s->taux[nd[3] + k];
                real t;
                if (hit_tri(s->tris + 12 * (size_t)i, o, d, tmin, *tmax, &t)) {
                    *found = i;
                    *tmax = t;
                    hit = 1;
                    if (any) return 1;
                }
            }
            continue;
        }
        if (sp + 2 > 512) abort();   /* deeper than any tree the tests build */
        stack[sp++] = nd[7];
        stack[sp++] = nd[3];
    }
    return hit;
}

/* every shape in id order, then every instance, the closest kept (intersect_world's brute
   force, accel.c:264-277); the triangles through the tree where the scene has one */
static int scene_intersect(const LroScene *s, V o, V d, real tmin, real tmax, Hit *h)
{
    int found = -1, finst = -1;
    int n = s->tnodes ? s->nprims : s->nprims + s->ntris;
    for (int i = 0; i < n; i++) {
        real t;
        if (shape_test(s, i, o, d, tmin, tmax, &t)) { found = i; tmax = t; }
    }
    if (s->tnodes && s->ntris > 0) {
        int ti;
        if (tree_tris(s, o, d, tmin, &tmax, 0, &ti)) found = s->nprims + ti;
    }
    for (int k = 0; k < s->ninst; k++) {
        real t;
        int local;
        if (inst_test(s, k, o, d, tmin, tmax, 0, &t, &local)) { found = local; finst = k; tmax = t; }
    }
    if (found < 0 && finst < 0) return 0;
    real t = tmax;
    V outward;
    if (finst >= 0) {
        const LroInst *in = &s->inst[finst];
        V oo, od;
        to_object(in, o, d, &oo, &od);
        V no;
        if (in->kind == LRO_INST_SPHERE) no = vadd(oo, vscale(od, t));
        else if (in->kind == LRO_INST_BOX) no = unit_box_normal(vadd(oo, vscale(od, t)));
        else {
            const float *tr = s->mtris + 12 * (size_t)(s->mesh_first[in->kind] + found);
            no = vcross(vload(tr + 4), vload(tr + 8));
        }
        outward = normal_to_world(in, no);
    } else if (found < s->nprims) {
        const LroPrim *p = &s->prims[found];
        if (p->q[3] < 0.5f) outward = vscale(vsub(vadd(o, vscale(d, t)), vload(p->c)), 1.0f / p->c[3]);
        else outward = vload(p->n);
    } else {
        const float *tr = s->tris + 12 * (size_t)(found - s->nprims);
        outward = vnorm(vcross(vload(tr + 4), vload(tr + 8)));
    }
    h->t = t;
    h->p = vadd(o, vscale(d, t));
    h->front = vdot(d, outward) < 0;
    h->n = h->front ? outward : vneg(outward);
    h->ng = h->n;
    h->prim = found;
    h->inst = finst;
    return 1;
}
static int scene_occluded(const LroScene *s, V o, V d, real tmin, real tmax)
{
    int n = s->tnodes ? s->nprims : s->nprims + s->ntris;
    for (int i = 0; i < n; i++) {
        real t;
        if (shape_test(s, i, o, d, tmin, tmax, &t)) return 1;
    }
    if (s->tnodes && s->ntris > 0) {
        int ti;
        real tm = tmax;
        if (tree_tris(s, o, d, tmin, &tm, 1, &ti)) return 1;
    }
    for (int k = 0; k < s->ninst; k++) {
        real t;
        int local;
        if (inst_test(s, k, o, d, tmin, tmax, 1, &t, &local)) return 1;
    }
    return 0;
}
static int shape_material(const LroScene *s, int id)
{
    if (id < s->nprims) return (int)s->prims[id].u[3];
    uint32_t w;
    memcpy(&w, s->tris + 12 * (size_t)(id - s->nprims) + 11, 4);
    return (int)(w >> 26);
}
/* an instance's material, else its triangle's own */
static int hit_material(const LroScene *s, const Hit *h)
{
    if (h->inst < 0) return shape_material(s, h->prim);
    const LroInst *in = &s->inst[h->inst];
    if (in->material >= 0 || in->kind < 0) return in->material >= 0 ? in->material : 0;
    uint32_t w;
    memcpy(&w, s->mtris + 12 * (size_t)(s->mesh_first[in->kind] + h->prim) + 11, 4);
    return (int)(w >> 26);
}

void lro_scene_trace(const LroScene *s, const float *rays, int n, int any, float *hits)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
        const float *r = rays + 8 * (size_t)i;
        V o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
        float *o4 = hits + 4 * (size_t)i;
        uint32_t none = 0xffffffffu, id, ins;
        if (any) {
            int hit = scene_occluded(s, o, d, r[3], r[7]);
            o4[0] = hit ? 0.0f : r[7];
            id = hit ? 0u : none;
            ins = none;
        } else {
            Hit h;
            if (scene_intersect(s, o, d, r[3], r[7], &h)) {
                o4[0] = h.t;
                /* a unit sphere or box instance: 0xfffffff0 / 0xfffffff1 (trace.wgsl:188-190) */
                id = h.prim >= 0 ? (uint32_t)h.prim : 0xfffffff0u + (uint32_t)(-1 - h.prim);
                ins = h.inst >= 0 ? (uint32_t)h.inst : none;
            } else {
                o4[0] = r[7];
                id = none;
                ins = none;
            }
        }
        memcpy(&o4[1], &id, 4);
        memcpy(&o4[2], &ins, 4);
        o4[3] = 0;
    }
}

/* ---- the sky (env.c) ---- */
static const float SKYPI = 3.14159265358979323846f;
static void env_uv(const LroEnv *e, V d, real *u, real *v)   /* env_uv + map_uv (env.c:36-57) */
{
    if (e->rot != 0) d = v3(d.x * e->rot_c - d.z * e->rot_s, d.y, d.x * e->rot_s + d.z * e->rot_c);
    real y = d.y < -1 ? -1 : d.y > 1 ? 1 : d.y;
    real uu = 0.5f + atan2f(d.x, d.z) * (0.5f / SKYPI);
    *u = uu - floorf(uu);
    *v = acosf(y) * (1.0f / SKYPI);
}
static V env_dir(const LroEnv *e, real u, real v)            /* env_dir + map_dir (env.c:48-68) */
{
    real ph = (u - 0.5f) * (2.0f * SKYPI), th = v * SKYPI;
    real st = sinf(th);
    V d = v3(st * sinf(ph), cosf(th), st * cosf(ph));
    return e->rot != 0 ? v3(d.x * e->rot_c + d.z * e->rot_s, d.y, -d.x * e->rot_s + d.z * e->rot_c) : d;
}
static V env_bilerp(const LroEnv *e, real u, real v)         /* bilerp (env.c:70-88) */
{
    const int w = e->w, h = e->h;
    real x = u * w - 0.5f, y = v * h - 0.5f;
    int x0 = (int)floorf(x), y0 = (int)floorf(y);
    real fx = x - x0, fy = y - y0;
    int y1 = y0 + 1;
    if (y0 < 0) y0 = 0;
    if (y1 > h - 1) y1 = h - 1;
    if (y0 > h - 1) y0 = h - 1;
    int xa = x0 < 0 ? x0 + w : (x0 >= w ? x0 - w : x0);
    if (xa < 0 || xa >= w) xa = ((xa % w) + w) % w;
    int xb = xa + 1 == w ? 0 : xa + 1;
    const float *a = e->rgb + 3 * ((size_t)y0 * w + xa), *b = e->rgb + 3 * ((size_t)y0 * w + xb);
    const float *c = e->rgb + 3 * ((size_t)y1 * w + xa), *dd = e->rgb + 3 * ((size_t)y1 * w + xb);
    return v3((1 - fy) * ((1 - fx) * a[0] + fx * b[0]) + fy * ((1 - fx) * c[0] + fx * dd[0]),
              (1 - fy) * ((1 - fx) * a[1] + fx * b[1]) + fy * ((1 - fx) * c[1] + fx * dd[1]),
              (1 - fy) * ((1 - fx) * a[2] + fx * b[2]) + fy * ((1 - fx) * c[2] + fx * dd[2]));
}
static V env_eval(const LroEnv *e, V d)
{
    real u, v;
    env_uv(e, d, &u, &v);
    return vscale(env_bilerp(e, u, v), e->scale);
}
static int env_upper_bound(const float *c, int n, real x)    /* the last i with c[i] <= x (env.c:105-112) */
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (c[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}
static V env_sample(const LroEnv *e, real u1, real u2, V *d, real *pdf)   /* env.c:171-189 */
{
    *pdf = 0;
    if (!(e->sum > 0)) return v3(0, 0, 0);
    int y = env_upper_bound(e->marg, e->h, u2);
    real m0 = e->marg[y], m1 = e->marg[y + 1];
    real dv = m1 > m0 ? (u2 - m0) / (m1 - m0) : 0.5f;
    const float *c = e->cond + (size_t)y * (e->w + 1);
    int x = env_upper_bound(c, e->w, u1);
    real c0 = c[x], c1 = c[x + 1];
    real du = c1 > c0 ? (u1 - c0) / (c1 - c0) : 0.5f;
    if (dv > 0.9999f) dv = 0.9999f;
    if (du > 0.9999f) du = 0.9999f;
    real u = (x + du) / e->w, v = (y + dv) / e->h;
    real st = sinf(v * SKYPI);
    if (st < 1e-6f) return v3(0, 0, 0);
    *d = env_dir(e, u, v);
    *pdf = e->prob[(size_t)y * e->w + x] * e->w * e->h / ((2.0f * SKYPI * SKYPI) * st);
    return env_eval(e, *d);
}
static real env_pdf(const LroEnv *e, V d)                    /* env.c:191-203 */
{
    if (!(e->sum > 0)) return 0;
    real u, v;
    env_uv(e, d, &u, &v);
    real st = sinf(v * SKYPI);
    if (st < 1e-6f) return 0;
    int x = (int)(u * e->w), y = (int)(v * e->h);
    if (x > e->w - 1) x = e->w - 1;
    if (y > e->h - 1) y = e->h - 1;
    return e->prob[(size_t)y * e->w + x] * e->w * e->h / ((2.0f * SKYPI * SKYPI) * st);
}

int lro_env_prepare(LroEnv *e, const float *rgb, int w, int h, float scale, float rot)
{
    memset(e, 0, sizeof(*e));
    e->w = w;
    e->h = h;
    e->rgb = rgb;
    e->scale = scale;
    e->rot = rot;
    e->rot_c = cosf(rot);
    e->rot_s = sinf(rot);
    e->marg = malloc(sizeof(float) * (size_t)(h + 1));
    e->cond = malloc(sizeof(float) * (size_t)h * (w + 1));
    e->prob = malloc(sizeof(float) * (size_t)w * h);
    float *lum = malloc(sizeof(float) * (size_t)w * h), *wm = malloc(sizeof(float) * (size_t)w * h);
    float *rs = malloc(sizeof(float) * (size_t)h);
    double *rows = malloc(sizeof(double) * (size_t)h);
    if (!e->marg || !e->cond || !e->prob || !lum || !wm || !rs || !rows) {
        free(lum); free(wm); free(rs); free(rows);
        lro_env_free(e);
        return -1;
    }
    /* make_weights (env.c:125-162) */
    for (int y = 0; y < h; y++) {
        const float *p = rgb + 3 * (size_t)y * w;
        float *l = lum + (size_t)y * w, *o = wm + (size_t)y * w;
        for (int x = 0; x < w; x++) l[x] = 0.2126f * p[3 * x] + 0.7152f * p[3 * x + 1] + 0.0722f * p[3 * x + 2];
        for (int x = 0; x < w; x++) {
            float a = l[x == 0 ? w - 1 : x - 1], b = l[x], c = l[x + 1 == w ? 0 : x + 1];
            float m = a > b ? a : b;
            o[x] = m > c ? m : c;
        }
        rs[y] = sinf((y + 0.5f) * SKYPI / h);
    }
    memcpy(lum, wm, sizeof(float) * (size_t)w * h);
    for (int y = 0; y < h; y++) {
        const float *a = lum + (size_t)(y > 0 ? y - 1 : y) * w, *b = lum + (size_t)y * w;
        const float *c = lum + (size_t)(y + 1 < h ? y + 1 : y) * w;
        float *o = wm + (size_t)y * w;
        for (int x = 0; x < w; x++) {
            float m = a[x] > b[x] ? a[x] : b[x];
            m = m > c[x] ? m : c[x];
            o[x] = m > 0 ? m : 0;
        }
    }
    /* the sampling tables of prepare_base (env.c:343-364) */
    e->sum = 0;
    for (int y = 0; y < h; y++) {
        float *c = e->cond + (size_t)y * (w + 1);
        double acc = 0;
        c[0] = 0;
        for (int x = 0; x < w; x++) {
            acc += wm[(size_t)y * w + x] * rs[y];
            c[x + 1] = (float)acc;
        }
        rows[y] = acc;
        for (int x = 1; x <= w; x++) c[x] = acc > 0 ? (float)(c[x] / acc) : (float)x / w;
        c[w] = 1;
        e->sum += acc;
    }
    double acc = 0;
    e->marg[0] = 0;
    for (int y = 0; y < h; y++) {
        acc += rows[y];
        e->marg[y + 1] = e->sum > 0 ? (float)(acc / e->sum) : (float)(y + 1) / h;
    }
    e->marg[h] = 1;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            e->prob[(size_t)y * w + x] = e->sum > 0 ? (float)((float)(wm[(size_t)y * w + x] * rs[y]) / e->sum) : 0.0f;
    free(lum); free(wm); free(rs); free(rows);
    return 0;
}
void lro_env_free(LroEnv *e)
{
    free(e->marg); free(e->cond); free(e->prob);
    e->marg = e->cond = e->prob = NULL;
}
void lro_env_eval(const LroEnv *e, const float d[3], float out[3])
{
    V c = env_eval(e, v3(d[0], d[1], d[2]));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
void lro_env_sample(const LroEnv *e, float u1, float u2, float d[3], float *pdf, float le[3])
{
    V dd = v3(0, 0, 0);
    real p;
    V c = env_sample(e, u1, u2, &dd, &p);
    d[0] = dd.x; d[1] = dd.y; d[2] = dd.z;
    *pdf = p;
    le[0] = c.x; le[1] = c.y; le[2] = c.z;
}
float lro_env_pdf(const LroEnv *e, const float d[3]) { return env_pdf(e, v3(d[0], d[1], d[2])); }

/* ---- next event estimation to the quad lamp (tracer.c:30-81) ---- */
static inline real power_heuristic(real pa, real pb)
{
    real a = pa * pa, b = pb * pb;
    return a + b > 0 ? a / (a + b) : 0;
}
static inline real light_pdf(const LroPrim *light, real dist2, real cos_light) { return dist2 / (cos_light * light->v[3]); }

static V sample_direct_light(const LroScene *s, const Hit *h, Bsdf *b, V v, int mis, real ua, real ub)
{
    const LroPrim *light = &s->prims[s->light];
    V le = vload(s->mats[(int)light->u[3]].emission);
    V xl = vadd(vload(light->q), vadd(vscale(vload(light->u), ua), vscale(vload(light->v), ub)));
    V to_light = vsub(xl, h->p);
    real dist2 = vdot(to_light, to_light);
    real dist = sqrtf(dist2);
    V wi = vscale(to_light, 1.0f / dist);
    real cos_surf = vdot(h->n, wi);
    real cos_light = -vdot(vload(light->n), wi);
    if (cos_light <= 0) return v3(0, 0, 0);
    real pdf;
    V f = bsdf_eval(b, h->n, v, wi, &pdf);
    if (!(vmax(f) > 0)) return v3(0, 0, 0);
    if (scene_occluded(s, h->p, wi, s->eps, dist - s->margin)) return v3(0, 0, 0);
    real g = fabsf(cos_surf) * cos_light * light->v[3] / dist2;
    if (mis) g *= power_heuristic(light_pdf(light, dist2, cos_light), pdf);
    return vscale(vmul(f, le), g);
}

static int rough_part(const Bsdf *b) { return !b->smooth || (1 - b->metallic) * (1 - b->transmission) > 0; }

/* next event estimation towards the sky (sample_direct_env, tracer.c:87-101) */
static V sample_direct_env(const LroScene *s, const Hit *h, Bsdf *b, V v, int mis, real ua, real ub)
{
    V wi;
    real pdf;
    V le = env_sample(s->env, ua, ub, &wi, &pdf);
    if (!(pdf > 0) || !(vmax(le) > 0)) return v3(0, 0, 0);
    real bpdf;
    V f = bsdf_eval(b, h->n, v, wi, &bpdf);
    if (!(vmax(f) > 0)) return v3(0, 0, 0);
    if (scene_occluded(s, h->p, wi, s->eps, 1e30f)) return v3(0, 0, 0);
    real g = fabsf(vdot(h->n, wi)) / pdf;
    if (mis) g *= power_heuristic(pdf, bpdf);
    return vscale(vmul(f, le), g);
}

/* ---- glowing shapes as lights (Scene.emitters: tracer.c:120-229, scene.c:1000-1031) ----
 * A glowing instance (a ball or a box) is sampled at rough vertices as the lamp is: one picked
 * uniformly, then a direction in the cone a ball fills (its density over solid angle), or a
 * point on a box's faces by area; a path that finds one after a rough bounce is weighted
 * against that (emit_pdf). The numbers are the plain stream's, three per vertex. */
static int glows(const LroMat *m)                                 /* material_glows, scene.h:87-89 */
{
    return m->type != LRO_MAT_LIGHT && (m->emission[0] > 0 || m->emission[1] > 0 || m->emission[2] > 0);
}

static V axis_col(const float *fwd, int k) { return v3(fwd[k], fwd[4 + k], fwd[8 + k]); }

int lro_find_emitters(const float *fwd, const int *kind, const int *material, int ninst, const LroMat *mats,
                      int nmats, LroEmitter *out, int *slot)
{
    int n = 0;
    for (int i = 0; i < ninst; i++) slot[i] = -1;
    for (int i = 0; i < ninst && n < LRO_MAX_EMITTERS; i++) {
        const float *x = fwd + 12 * (size_t)i;
        if ((kind[i] != LRO_INST_SPHERE && kind[i] != LRO_INST_BOX) || material[i] < 0 || material[i] >= nmats) continue;
        if (!glows(&mats[material[i]])) continue;
        V a[3] = {axis_col(x, 0), axis_col(x, 1), axis_col(x, 2)};
        const int ball = kind[i] == LRO_INST_SPHERE;
        if (ball) {   /* a stretched ball is no light */
            real la = vlen(a[0]), lb = vlen(a[1]), lc = vlen(a[2]);
            if (!(la > 0) || fabsf(la - lb) > 1e-4f * la || fabsf(la - lc) > 1e-4f * la) continue;
        }
        LroEmitter *e = &out[n];
        memset(e, 0, sizeof(*e));
        e->ball = ball;
        e->c[0] = x[3];
        e->c[1] = x[7];
        e->c[2] = x[11];
        for (int k = 0; k < 3; k++) {
            e->a[3 * k] = (float)a[k].x;
            e->a[3 * k + 1] = (float)a[k].y;
            e->a[3 * k + 2] = (float)a[k].z;
        }
        for (int k = 0; k < 3; k++) e->le[k] = mats[material[i]].emission[k];
        if (ball) {
            real r = vlen(a[0]);
            e->r = (float)r;
            e->area = (float)(4 * (real)PI_ * r * r);
        } else
            e->area = (float)(8 * (vlen(vcross(a[1], a[2])) + vlen(vcross(a[0], a[2])) + vlen(vcross(a[0], a[1]))));
        slot[i] = n++;
    }
    return n;
}

/* 1 - cos of the half angle of the cone a ball of radius r fills from dist2 (squared) away, in
 * the form without cancellation for a small far ball (ball_cone, tracer.c:140-146); 0 inside */
static real ball_cone(real r, real dist2)
{
    real sin2 = r * r / dist2;
    if (!(sin2 < 1)) return 0;
    return sin2 / (1 + sqrtf(1 - sin2));
}

static V emitter_axis(const LroEmitter *e, int k) { return vload(e->a + 3 * k); }

/* emit_pdf (tracer.c:148-158): the density over solid angle, the pick included, that NEE would
 * have given the point of emitter k at distance t along unit d from p, its outward normal n */
static real emit_pdf(const LroScene *s, int k, V p, V d, real t, V n)
{
    const LroEmitter *e = &s->emit[k];
    if (e->ball) {
        V to = vsub(vload(e->c), p);
        real om = ball_cone(e->r, vdot(to, to));
        return om > 0 ? 1 / (2 * (real)PI_ * om * s->nemit) : 0;
    }
    real cl = -vdot(n, d);
    return cl > 0 ? t * t / (cl * e->area * s->nemit) : 0;
}

/* emit_sample (tracer.c:160-200): a point of a glowing shape seen from p */
static int emit_sample(const LroScene *s, V p, real u0, real ua, real ub, V *wi, real *dist, real *pdf, V *le)
{
    int k = (int)(u0 * s->nemit);
    if (k >= s->nemit) k = s->nemit - 1;
    const LroEmitter *e = &s->emit[k];
    *le = vload(e->le);
    if (e->ball) {
        V to = vsub(vload(e->c), p);
        real d2 = vdot(to, to), om = ball_cone(e->r, d2);
        if (!(om > 0)) return 0;
        real dc = sqrtf(d2), cos_t = 1 - ua * om, sin_t = sqrtf(rfmax(0, 1 - cos_t * cos_t));
        real ph = 2 * (real)PI_ * ub;
        V w = vscale(to, 1 / dc), t, b;
        onb(w, &t, &b);
        *wi = vadd(vscale(w, cos_t), vadd(vscale(t, sin_t * cosf(ph)), vscale(b, sin_t * sinf(ph))));
        real disc = e->r * e->r - d2 * sin_t * sin_t;   /* the ball's near side along wi */
        *dist = dc * cos_t - sqrtf(rfmax(0, disc));
        *pdf = 1 / (2 * (real)PI_ * om * s->nemit);
        return *dist > 0;
    }
    /* a box: one of its six faces (two per axis, 2 a_j by 2 a_k) by area, then a point on it */
    V a[3] = {emitter_axis(e, 0), emitter_axis(e, 1), emitter_axis(e, 2)};
    real fa[3];
    for (int i = 0; i < 3; i++) fa[i] = 4 * vlen(vcross(a[(i + 1) % 3], a[(i + 2) % 3]));
    real x = ua * e->area;
    int f = 0;
    while (f < 5 && x >= fa[f >> 1]) {
        x -= fa[f >> 1];
        f++;
    }
    int i = f >> 1;
    real sg = f & 1 ? -1 : 1, uu = rfmin(x / fa[i], 1) * 2 - 1, vv = ub * 2 - 1;
    V aj = a[(i + 1) % 3], ak = a[(i + 2) % 3];
    V q = vadd(vload(e->c), vadd(vscale(a[i], sg), vadd(vscale(aj, uu), vscale(ak, vv))));
    V n = vscale(vnorm(vcross(aj, ak)), sg);
    if (vdot(n, a[i]) * sg < 0) n = vneg(n);            /* outward, whatever the axes' handedness */
    V to = vsub(q, p);
    real d2 = vdot(to, to);
    *dist = sqrtf(d2);
    *wi = vscale(to, 1 / *dist);
    real cl = -vdot(n, *wi);
    if (!(cl > 0)) return 0;
    *pdf = d2 / (cl * e->area * s->nemit);
    return 1;
}

/* sample_direct_emit (tracer.c:202-216): NEE of the glowing shapes at a surface */
static V sample_direct_emit(const LroScene *s, const Hit *h, Bsdf *b, V v, int mis, LroRng *rng)
{
    real u0 = lro_rng_float(rng), ua = lro_rng_float(rng), ub = lro_rng_float(rng);
    V wi, le;
    real dist, pdf;
    if (!emit_sample(s, h->p, u0, ua, ub, &wi, &dist, &pdf, &le)) return v3(0, 0, 0);
    real bpdf;
    V f = bsdf_eval(b, h->n, v, wi, &bpdf);
    if (!(vmax(f) > 0)) return v3(0, 0, 0);
    if (scene_occluded(s, h->p, wi, s->eps, dist - s->margin)) return v3(0, 0, 0);
    real g = fabsf(vdot(h->n, wi)) / pdf;
    if (mis) g *= power_heuristic(pdf, bpdf);
    return vscale(vmul(f, le), g);
}

enum { EMIT_FULL, EMIT_NONE, EMIT_MIS };

/* trace_path_sampled (tracer.c:364-509) without media or textures */
static V trace_path(const LroScene *s, V o, V d, Smp *sm, int max_depth, int mis, int *cam_glow)
{
    *cam_glow = 0;
    V radiance = v3(0, 0, 0), throughput = v3(1, 1, 1);
    int emit = EMIT_FULL;
    real bounce_pdf = 0;
    /* skip_caustics (tracer.c:371-375): the camera saw a rough surface first (chain) and the
       last vertex was a mirror or glass (spec): light tracing has that lamp hit */
    int caustic_chain = 0, caustic_spec = 0;
    if (mis) d = vnorm(d);
    for (int depth = 0;; depth++) {
        Hit h;
        if (!scene_intersect(s, o, d, s->eps, 1e30f, &h)) {   /* escaped: the sky (tracer.c:415-421) */
            if (s->env && emit != EMIT_NONE) {
                V le = env_eval(s->env, d);
                if (emit == EMIT_MIS) le = vscale(le, power_heuristic(bounce_pdf, env_pdf(s->env, d)));
                radiance = vadd(radiance, vmul(throughput, le));
            }
            break;
        }
        const LroMat *m = &s->mats[hit_material(s, &h)];
        if (m->type == LRO_MAT_LIGHT) {
            if (h.front && !(caustic_chain && caustic_spec)) {
                if (emit == EMIT_FULL) radiance = vadd(radiance, vmul(throughput, vload(m->emission)));
                else if (emit == EMIT_MIS && h.prim == s->light) {
                    const LroPrim *l = &s->prims[s->light];
                    real cos_light = -vdot(vload(l->n), d);
                    if (cos_light > 0) {
                        real w = power_heuristic(bounce_pdf, light_pdf(l, h.t * h.t, cos_light));
                        radiance = vadd(radiance, vmul(throughput, vscale(vload(m->emission), w)));
                    }
                }
            }
            break;
        }
        /* the camera ray found a surface glowing at least as bright as 1 (not a lamp): the
           denoiser's moments leave the sample out (trace.wgsl:1183-1186, 1309) */
        if (depth == 0 && rfmax(m->emission[0], rfmax(m->emission[1], m->emission[2])) >= 1) *cam_glow = 1;
        /* a glowing surface (tracer.c:450-458): found by the path; a glowing shape after a rough
           bounce was sampled by NEE: weighted against it (MIS) or counted there already */
        if (glows(m)) {
            int k = s->nemit && emit != EMIT_FULL && h.inst >= 0 ? s->inst_emit[h.inst] : -1;
            real pe = k >= 0 && h.front ? emit_pdf(s, k, o, vnorm(d), h.t * vlen(d), h.ng) : 0;
            if (!(pe > 0)) radiance = vadd(radiance, vmul(throughput, vload(m->emission)));
            else if (emit == EMIT_MIS)
                radiance = vadd(radiance, vmul(throughput, vscale(vload(m->emission), power_heuristic(bounce_pdf, pe))));
        }
        if (depth >= max_depth) break;
        V v = vneg(vnorm(d));
        Bsdf b = bsdf_make(m, h.front, s->tab);
        if (rough_part(&b)) {
            real la, lb;
            if (s->light >= 0) {
                sampler_2d(sm, LRO_SAMPLE_LIGHT(depth), &la, &lb);
                radiance = vadd(radiance, vmul(throughput, sample_direct_light(s, &h, &b, v, mis, lb, la)));
            }
            if (s->env) {   /* env_numbers (tracer.c:232-235) */
                if (s->light < 0) sampler_2d(sm, LRO_SAMPLE_LIGHT(depth), &la, &lb);
                else {
                    la = lro_rng_float(sm->rng);
                    lb = lro_rng_float(sm->rng);
                }
                radiance = vadd(radiance, vmul(throughput, sample_direct_env(s, &h, &b, v, mis, la, lb)));
            }
            if (s->nemit) radiance = vadd(radiance, vmul(throughput, sample_direct_emit(s, &h, &b, v, mis, sm->rng)));
        }
        bsdf_view(&b, vdot(v, h.n));
        real ur = 0, ul = 0, da = 0, db = 0;
        if (depth >= 3 || bsdf_nlobes(&b) > 1) sampler_2d(sm, LRO_SAMPLE_MISC(depth), &ur, &ul);
        int lobe = bsdf_pick(&b, ul);
        if (lobe >= 0 && !bsdf_lobe_delta(&b, lobe)) sampler_2d(sm, LRO_SAMPLE_DIR(depth), &da, &db);
        BSample bs;
        if (!bsdf_sample_lobe(&b, lobe, h.n, v, da, db, &bs)) break;
        throughput = vmul(throughput, bs.w);
        if (bs.delta) {
            emit = EMIT_FULL;
            caustic_spec = 1;
        } else {
            bounce_pdf = bs.pdf;
            emit = mis ? EMIT_MIS : EMIT_NONE;
            if (depth == 0) caustic_chain = s->skip_caustics;
            caustic_spec = 0;
        }
        if (depth >= 3) {                                    /* russian_roulette_u (tracer.c:295-301) */
            real pr = rfmin(vmax(throughput), 0.95f);
            if (ur >= pr) break;
            throughput = vscale(throughput, 1.0f / pr);
        }
        o = h.p;
        d = mis ? vnorm(bs.l) : bs.l;
    }
    return radiance;
}

/* the engine's fixed point: a sample's radiance in units of 2^-28, saturated at 2^24, rounded to
   the nearest (even) */
static long long to_fixed(real x)
{
    x = x > 16777216.0f ? 16777216.0f : (x < -16777216.0f ? -16777216.0f : x);
    return llrint((double)(x * (real)268435456.0f));
}

/* ---- the photon cache (deposit / pc_claim, trace.wgsl:1359-1404; trace_decay.wgsl; the lookup of
 * trace_resolve.wgsl:43-80), as the library lays it out: LRO_PC_SLOTS slots of 4 x 64 bits (rgb
 * sums in 2^-8 units; the photons in 1/16 units in the low 32 bits of the fourth word, the cell's
 * tag in the high 32). A cell's key (its integer corner, 19 bits an axis, the axis of the
 * normal and its sign 3 bits, the size 2 bits) mixed by SplitMix64's finaliser (Steele, Lea and
 * Flood 2014) gives its home slot (low bits) and its 32-bit tag (high bits, never 0); a cell
 * takes the first slot from its home on that is free or its own, at most 8 on. Here one path
 * after another: which of two cells meeting in a slot gets it can differ from the device's
 * (whose order is the hardware's), never a cell's sums. */
static void camera_ray(const float *cam, real s, real t, V *o, V *d);
static uint64_t pc_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
uint64_t lro_pc_key(int x, int y, int z, int axis, int level)
{
    const uint64_t m = (1ULL << 19) - 1ULL;
    return pc_mix(((uint64_t)(uint32_t)x & m) | (((uint64_t)(uint32_t)y & m) << 19) | (((uint64_t)(uint32_t)z & m) << 38) |
                  ((uint64_t)axis << 57) | ((uint64_t)level << 60));
}
static int pc_axis(V n)
{
    const real ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    const int a = ax >= ay && ax >= az ? 0 : (ay >= az ? 1 : 2);
    const real v = a == 0 ? n.x : (a == 1 ? n.y : n.z);
    return 2 * a + (v < 0 ? 1 : 0);
}
static uint32_t pc_tag(uint64_t h) { return (uint32_t)(h >> 32) | 1u; }
static uint32_t pc_claim(uint64_t *t, uint64_t h)
{
    const uint32_t tag = pc_tag(h);
    uint32_t idx = (uint32_t)h & (LRO_PC_SLOTS - 1u);
    for (int probe = 0; probe < 8; probe++) {
        const uint32_t k = (uint32_t)(t[4 * (size_t)idx + 3] >> 32);
        if (k == 0u) t[4 * (size_t)idx + 3] |= (uint64_t)tag << 32;
        if (k == 0u || k == tag) return idx;
        idx = (idx + 1u) & (LRO_PC_SLOTS - 1u);
    }
    return 0xffffffffu;
}
/* a point's cell: along the normal's main axis offset by 1/pi of a cell (a surface at a round
 * coordinate then lies inside a layer of cells, not on the boundary of two) */
static int pc_coord(float x, float size, int normal_axis)
{
    return (int)floorf(x / size + (normal_axis ? 0.31830988f : 0.0f));
}
/* a photon of power c at p on a surface facing n, into its cell of each size (cell, x4, x16) */
static void pc_deposit(uint64_t *t, float cell, V p, V n, V c)
{
    const int axis = pc_axis(n), ax = axis / 2;
    float size = cell;
    for (int level = 0; level < 3; level++, size *= 4.0f) {
        const uint32_t slot = pc_claim(t, lro_pc_key(pc_coord((float)p.x, size, ax == 0), pc_coord((float)p.y, size, ax == 1),
                                                     pc_coord((float)p.z, size, ax == 2), axis, level));
        if (slot == 0xffffffffu) continue;
        uint64_t *w = t + 4 * (size_t)slot;
        const float cc[3] = {(float)c.x, (float)c.y, (float)c.z};
        for (int ch = 0; ch < 3; ch++) {
            const float v = cc[ch] * 256.0f;
            if (!(v > 0) || v > 1.8e18f) continue;
            w[ch] += (uint64_t)llrintf(v);
        }
        w[3] += 16u;
    }
}
void lro_pc_deposit(uint64_t *table, float cell, const float p[3], const float n[3], const float c[3])
{
    pc_deposit(table, cell, v3(p[0], p[1], p[2]), v3(n[0], n[1], n[2]), v3(c[0], c[1], c[2]));
}
void lro_pc_deposit_many(uint64_t *table, float cell, int count, const float *p, const float *n, const float *c)
{
    for (int i = 0; i < count; i++) lro_pc_deposit(table, cell, p + 3 * (size_t)i, n + 3 * (size_t)i, c + 3 * (size_t)i);
}
/* the fade: sums and photons times k; a cell whose photons faded to nothing gives its slot back
 * when the next slot is free (as it was before this fade) */
void lro_pc_fade(uint64_t *t, double k)
{
    uint32_t first_tag = (uint32_t)(t[3] >> 32);
    for (uint32_t s = 0; s < LRO_PC_SLOTS; s++) {
        uint64_t *w = t + 4 * (size_t)s;
        const uint64_t w3 = w[3];
        if ((uint32_t)(w3 >> 32) == 0u) continue;
        for (int ch = 0; ch < 3; ch++) w[ch] = (uint64_t)llrint((double)w[ch] * k);
        const uint32_t count = (uint32_t)floor((double)(uint32_t)w3 * k);
        const uint32_t next = s + 1u < LRO_PC_SLOTS ? (uint32_t)(t[4 * (size_t)(s + 1u) + 3] >> 32) : first_tag;
        if (count == 0u && next == 0u) w[0] = w[1] = w[2] = w[3] = 0;
        else w[3] = (w3 & 0xffffffff00000000ULL) | count;
    }
}
static void pc_sum(const uint64_t *t, uint64_t h, double e[4])
{
    const uint32_t tag = pc_tag(h);
    uint32_t idx = (uint32_t)h & (LRO_PC_SLOTS - 1u);
    e[0] = e[1] = e[2] = e[3] = 0;
    for (int probe = 0; probe < 8; probe++) {
        const uint64_t *w = t + 4 * (size_t)idx;
        const uint32_t k = (uint32_t)(w[3] >> 32);
        if (k == tag) {
            e[0] = (double)w[0];
            e[1] = (double)w[1];
            e[2] = (double)w[2];
            e[3] = (double)(uint32_t)w[3] / 16.0;
            return;
        }
        if (k == 0u) return;
        idx = (idx + 1u) & (LRO_PC_SLOTS - 1u);
    }
}
/* the lookup (trace_resolve.wgsl:43-80) at each pixel's centre ray: where it first meets a rough
 * surface, albedo / pi x the power of the finest cells around the point (bilinearly, in the
 * plane of its normal's main axis) holding at least 16 photons (else the largest), over their
 * area and the light paths (x 256); out: 4 floats a pixel (rgb, 1; 0 where no rough surface) */
void lro_pc_resolve(const LroScene *s, const uint64_t *t, float cell, double paths, const float cam[12], int W, int H, float *out)
{
    const double norm = paths > 0 ? 1.0 / (paths * 256.0) : 0.0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int p = 0; p < W * H; p++) {
        const int x = p % W, y = p / W;
        V o, d;
        camera_ray(cam, ((real)x + 0.5f) / W, ((real)y + 0.5f) / H, &o, &d);
        float *r = out + 4 * (size_t)p;
        r[0] = r[1] = r[2] = r[3] = 0;
        Hit h;
        if (!scene_intersect(s, o, d, s->eps, 1e30f, &h)) continue;
        const LroMat *m = &s->mats[hit_material(s, &h)];
        if (m->type == LRO_MAT_LIGHT) continue;
        Bsdf b = bsdf_make(m, h.front, s->tab);
        if (!rough_part(&b)) continue;
        const int axis = pc_axis(h.n), ax = axis / 2, t1 = (ax + 1) % 3, t2 = (ax + 2) % 3;
        const float pc3[3] = {(float)h.p.x, (float)h.p.y, (float)h.p.z};
        float size = cell;
        double e[4] = {0, 0, 0, 0};
        for (int level = 0; level < 3; level++) {
            float fp[3];
            int c0[3];
            for (int k = 0; k < 3; k++) {
                fp[k] = pc3[k] / size;
                c0[k] = pc_coord(pc3[k], size, k == ax);
            }
            const float uu = fp[t1] - 0.5f, vv = fp[t2] - 0.5f;
            const float u0 = floorf(uu), v0 = floorf(vv), fu = uu - u0, fv = vv - v0;
            e[0] = e[1] = e[2] = e[3] = 0;
            for (int j = 0; j < 4; j++) {
                int c[3] = {c0[0], c0[1], c0[2]};
                c[t1] = (int)u0 + (j & 1);
                c[t2] = (int)v0 + (j >> 1);
                const double w = ((j & 1) ? fu : 1.0f - fu) * ((j >> 1) ? fv : 1.0f - fv);
                double cs[4];
                pc_sum(t, lro_pc_key(c[0], c[1], c[2], axis, level), cs);
                for (int q = 0; q < 4; q++) e[q] += w * cs[q];
            }
            if (e[3] >= 16.0 || level == 2) break;
            size *= 4.0f;
        }
        const double k = norm / ((double)size * size) / 3.14159265358979323846;
        r[0] = (float)(m->albedo[0] * e[0] * k);
        r[1] = (float)(m->albedo[1] * e[1] * k);
        r[2] = (float)(m->albedo[2] * e[2] * k);
        r[3] = 1;
    }
}

/* light_trace_caustics (tracer.c:511-665) as the engine issues it: npass passes of nlp light
 * paths each (pass k, path i: the PCG32 stream lro_pt_seed(i, 2^31 + s0 + k, seed)). A path
 * starts at a uniform point of the lamp (two numbers) in a cosine-distributed direction about
 * its normal (two more), power Le area pi; the first vertex must be a mirror or glass; at every
 * rough vertex after it the path is joined to the camera (camera_project, camera.h:48-59) by a
 * shadow ray and c = beta f |cos_o| / (dist^2 pixel_area cos_c^3) / nlp is added to the pixel it
 * lands in, each channel in the fixed point above (atomic integer sums: any order gives the same
 * acc, H x W x 3); then on by the BSDF seen from the light's side (three numbers), through a
 * surface times eta^2, Russian roulette from depth 3 (one more number). One normal per hit: the
 * adjoint correction of smooth normals is 1. No media. */
static void light_paths(const LroScene *s, const float cam[12], int W, int H, uint32_t nlp, uint32_t s0, int npass,
                        uint64_t seed, int max_depth, long long *acc, uint64_t *pc, float pc_cell);
void lro_light_trace(const LroScene *s, const float cam[12], int W, int H, uint32_t nlp, uint32_t s0, int npass,
                     uint64_t seed, int max_depth, long long *acc)
{
    light_paths(s, cam, W, H, nlp, s0, npass, seed, max_depth, acc, NULL, 0);
}
/* the cache mode (lrh_photons_trace): npaths light paths (sample index 2^31 + s0), each leaving
 * its power at every rough vertex after the first in the table instead of joining the camera;
 * one path after another */
void lro_photons_trace(const LroScene *s, uint64_t *table, float cell, uint32_t npaths, uint32_t s0, uint64_t seed,
                       int max_depth)
{
    static const float nocam[12] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 1, 0};
    light_paths(s, nocam, 1, 1, npaths, s0, 1, seed, max_depth, NULL, table, cell);
}
static void light_paths(const LroScene *s, const float cam[12], int W, int H, uint32_t nlp, uint32_t s0, int npass,
                        uint64_t seed, int max_depth, long long *acc, uint64_t *pc, float pc_cell)
{
    if (s->light < 0 || nlp == 0 || npass <= 0) return;
    const LroPrim *lamp = &s->prims[s->light];
    const V le = vload(s->mats[(int)lamp->u[3]].emission);
    const V org = vload(cam), ul = vload(cam + 3), hz = vload(cam + 6), vt = vload(cam + 9);
    const real pix_area = vlen(hz) * vlen(vt) / ((real)W * (real)H);
    const V center = vadd(ul, vsub(vscale(hz, 0.5f), vscale(vt, 0.5f)));
    const V fw = vsub(center, org);
    const real inv_n = 1.0f / (real)nlp;
    const long long total = (long long)nlp * npass;
#pragma omp parallel for schedule(dynamic, 1024) if (!pc)
    for (long long g = 0; g < total; g++) {
        const uint32_t pass = (uint32_t)(g / nlp), path = (uint32_t)(g % nlp);
        LroRng rng;
        lro_pt_seed(&rng, path, 0x80000000u + s0 + pass, seed);
        const real e1 = lro_rng_float(&rng), e2 = lro_rng_float(&rng);
        V o = vadd(vload(lamp->q), vadd(vscale(vload(lamp->u), e1), vscale(vload(lamp->v), e2)));
        const real c1 = lro_rng_float(&rng), c2 = lro_rng_float(&rng);
        V d = cosine_hemisphere(vload(lamp->n), c1, c2);
        V beta = vscale(le, lamp->v[3] * (real)3.14159265358979323846), thr = v3(1, 1, 1);
        for (int depth = 0; depth < max_depth; depth++) {
            Hit h;
            if (!scene_intersect(s, o, d, s->eps, 1e30f, &h)) break;
            const LroMat *m = &s->mats[hit_material(s, &h)];
            if (m->type == LRO_MAT_LIGHT) break;
            Bsdf b = bsdf_make(m, h.front, s->tab);
            const V wi = vneg(d);
            /* 1. the rough part joined to the camera, or left in the photon cache (lamp -> rough at
               the first vertex: NEE's) */
            if (pc && depth > 0 && rough_part(&b)) pc_deposit(pc, pc_cell, h.p, h.n, beta);
            if (!pc && depth > 0 && rough_part(&b)) {
                const V to_cam = vsub(org, h.p);
                const real dist2 = vdot(to_cam, to_cam), dist = sqrtf(dist2);
                const V wo = vscale(to_cam, 1.0f / dist);
                const real cos_o = vdot(h.n, wo);
                /* the BSDF seen from the camera: its side's normal and index */
                Bsdf bc = cos_o > 0 ? b : bsdf_make(m, !h.front, s->tab);
                const V nc = cos_o > 0 ? h.n : vneg(h.n);
                real pdf;
                V f = v3(0, 0, 0);
                if (cos_o != 0 && vdot(nc, wo) > 0) f = bsdf_eval(&bc, nc, wo, wi, &pdf);
                const V qv = vsub(h.p, org);
                const real z = vdot(qv, fw);
                if (vmax(f) > 0 && z > 0) {
                    const V e = vsub(vscale(qv, 1.0f / z), fw);
                    const real sx = 0.5f + vdot(e, hz) / vdot(hz, hz), sy = 0.5f - vdot(e, vt) / vdot(vt, vt);
                    const real cos_c = z / vlen(qv);
                    if (sx >= 0 && sy >= 0 && sx < 1 && sy < 1) {
                        const int px = (int)(sx * (real)W), py = (int)(sy * (real)H);
                        if (px < W && py < H && !scene_occluded(s, h.p, wo, s->eps, dist * (1.0f - 1e-4f))) {
                            const real wgt = fabsf(cos_o) / (dist2 * pix_area * cos_c * cos_c * cos_c);
                            const V c = vscale(vmul(beta, f), wgt * inv_n);
                            const real cc[3] = {c.x, c.y, c.z};
                            long long *a = acc + 3 * ((size_t)py * W + px);
                            for (int ch = 0; ch < 3; ch++) {
                                if (!(cc[ch] > 0)) continue;
                                const long long v = to_fixed(cc[ch]);
#pragma omp atomic
                                a[ch] += v;
                            }
                        }
                    }
                }
            }
            /* 2. on, by the BSDF seen from the light's side (bsdf_sample, brdf.h:343-348) */
            const real ulb = lro_rng_float(&rng), u1 = lro_rng_float(&rng), u2 = lro_rng_float(&rng);
            const real cv = vdot(wi, h.n);
            if (!(cv > 0)) break;
            bsdf_view(&b, cv);
            BSample bs;
            if (!bsdf_sample_lobe(&b, bsdf_pick(&b, ulb), h.n, wi, u1, u2, &bs)) break;
            if (depth == 0 && !bs.delta) break;
            const real og = vdot(h.n, bs.l);
            const int through = og < 0;   /* the sample's own flag, with one normal a hit */
            if (!bs.delta && (through ? og >= 0 : og <= 0)) break;
            V f = bs.w;
            if (through) f = vscale(f, b.eta * b.eta);
            beta = vmul(beta, f);
            thr = vmul(thr, f);
            if (depth >= 3) {
                const real pr = rfmin(vmax(thr), 0.95f);
                if (lro_rng_float(&rng) >= pr) break;
                beta = vscale(beta, 1.0f / pr);
                thr = vscale(thr, 1.0f / pr);
            }
            o = h.p;
            d = vnorm(bs.l);
        }
    }
}

/* camera_ray (camera.h:38-42) */
static void camera_ray(const float *cam, real s, real t, V *o, V *d)
{
    V target = vsub(vadd(vload(cam + 3), vscale(vload(cam + 6), s)), vscale(vload(cam + 9), t));
    *o = vload(cam);
    *d = vnorm(vsub(target, *o));
}

void lro_pt_seed(LroRng *rng, uint64_t pixel, uint32_t k, uint64_t seed)
{
    lro_rng_seed(rng, 0x853c49e6748fea9bULL ^ (pixel * 0x9E3779B97F4A7C15ULL) ^ (seed * 0xD1B54A32D192ED03ULL) ^
                          ((uint64_t)k * 0xA0761D6478BD642FULL), pixel);
}

/* The denoiser's guide (trace.wgsl:1034-1107, guide): the ray through each pixel's centre,
 * through mirrors and glass (refracted unless the glass reflects all) to the first surface light
 * scatters on (kind 0: its normal facing the ray, its distance along the way, the albedo of the
 * way); a lamp or a surface glowing at least as bright as 1 (glow_bright, :1115): its normal and
 * distance, kind 1, albedo 1; nothing within 6 vertices: kind 1, distance 1e30. 8 floats a
 * pixel: n.xyz, z | albedo.rgb, kind. */
void lro_guide(const LroScene *s, const float cam[12], int W, int H, float *out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            V o, d;
            camera_ray(cam, ((real)x + 0.5f) / W, ((real)y + 0.5f) / H, &o, &d);
            V gn = v3(0, 0, 0), ga = v3(1, 1, 1), a = v3(1, 1, 1);
            real gz = 1e30f, kind = 1, dist = 0;
            for (int i = 0; i < 6; i++) {
                Hit h;
                if (!scene_intersect(s, o, d, s->eps, 1e30f, &h)) break;
                const LroMat *m = &s->mats[hit_material(s, &h)];
                Bsdf b = bsdf_make(m, h.front, s->tab);
                if (m->type == LRO_MAT_LIGHT || rfmax(m->emission[0], rfmax(m->emission[1], m->emission[2])) >= 1) {
                    gn = h.n;
                    gz = dist + h.t;
                    break;
                }
                if (rough_part(&b)) {
                    gn = h.n;
                    gz = dist + h.t;
                    ga = vmul(a, vload(m->albedo));
                    kind = 0;
                    break;
                }
                a = vmul(a, vload(m->albedo));
                dist += h.t;
                V nxt = vreflect(d, h.n);
                if (b.transmission > 0.5f && b.metallic < 0.5f) {
                    real cos_i = rfmin(vdot(vneg(d), h.n), 1);
                    if (b.eta * b.eta * (1 - cos_i * cos_i) <= 1) nxt = vrefract(d, h.n, b.eta);
                }
                o = h.p;
                d = vnorm(nxt);
            }
            float *g = out + 8 * ((size_t)y * W + x);
            g[0] = (float)gn.x; g[1] = (float)gn.y; g[2] = (float)gn.z; g[3] = (float)gz;
            g[4] = (float)ga.x; g[5] = (float)ga.y; g[6] = (float)ga.z; g[7] = (float)kind;
        }
}

static void pt_window(const LroScene *s, const float cam[12], int W, int H, int x0, int y0, int cw, int ch, uint32_t s0,
                      int ns, uint64_t seed, float clamp, int max_depth, int mis, float *sum, const float *guide,
                      float *moments);

/* trace_image (tracer.c:329-356) over the pixels x0 .. x0+cw-1, y0 .. y0+ch-1 of a W x H image:
 * the sum of samples s0 .. s0+ns-1 of each into sum (ch x cw, rgb, float, the samples added in
 * order; NaNs skipped, clamp applied) */
void lro_pt_window(const LroScene *s, const float cam[12], int W, int H, int x0, int y0, int cw, int ch, uint32_t s0,
                   int ns, uint64_t seed, float clamp, int max_depth, int mis, float *sum)
{
    pt_window(s, cam, W, H, x0, y0, cw, ch, s0, ns, seed, clamp, max_depth, mis, sum, NULL, NULL);
}

/* the whole image, and the denoiser's moments beside (trace.wgsl:1301-1308): per pixel the sums
 * of its samples' irradiance luminance (the sample over the guide's albedo on a surface) and of
 * its square (2 floats, added to moments) */
void lro_pt_image_moments(const LroScene *s, const float cam[12], int W, int H, uint32_t s0, int ns, uint64_t seed,
                          float clamp, int max_depth, int mis, const float *guide, float *sum, float *moments)
{
    pt_window(s, cam, W, H, 0, 0, W, H, s0, ns, seed, clamp, max_depth, mis, sum, guide, moments);
}

static void pt_window(const LroScene *s, const float cam[12], int W, int H, int x0, int y0, int cw, int ch, uint32_t s0,
                      int ns, uint64_t seed, float clamp, int max_depth, int mis, float *sum, const float *guide,
                      float *moments)
{
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int y = y0; y < y0 + ch; y++)
        for (int x = x0; x < x0 + cw; x++) {
            uint64_t pixel = (uint64_t)y * W + x;
            V acc = v3(0, 0, 0);
            double m1 = 0, m2 = 0;
            for (int q = 0; q < ns; q++) {
                uint32_t k = s0 + (uint32_t)q;
                LroRng rng;
                lro_pt_seed(&rng, pixel, k, seed);
                Smp sm = {&rng, k, (uint32_t)seed * 0x9e3779b9u, x, y, 1, s->mask};
                real jx, jy;
                sampler_2d(&sm, 0, &jx, &jy);
                real u = (x + jx) / W, v = (y + jy) / H;
                V o, d;
                camera_ray(cam, u, v, &o, &d);
                int cam_glow;
                V c = trace_path(s, o, d, &sm, max_depth, mis, &cam_glow);
                if (!(c.x == c.x && c.y == c.y && c.z == c.z)) continue;
                if (clamp > 0 && vmax(c) > clamp) c = vscale(c, clamp / vmax(c));
                acc = vadd(acc, c);
                if (moments && !(cam_glow && guide[8 * pixel + 7] == 0)) {
                    const float *ga = guide + 8 * pixel + 4;
                    V ir = ga[3] == 0 ? v3(c.x / rfmax(ga[0], 1e-3f), c.y / rfmax(ga[1], 1e-3f), c.z / rfmax(ga[2], 1e-3f)) : c;
                    float l = (float)lum(ir);
                    m1 += l;
                    m2 += (double)(l * l);
                }
            }
            float *o3 = sum + 3 * ((size_t)(y - y0) * cw + (x - x0));
            o3[0] += acc.x; o3[1] += acc.y; o3[2] += acc.z;
            if (moments) {
                moments[2 * pixel] += (float)m1;
                moments[2 * pixel + 1] += (float)m2;
            }
        }
}

/* the whole image */
void lro_pt_image(const LroScene *s, const float cam[12], int W, int H, uint32_t s0, int ns, uint64_t seed, float clamp,
                  int max_depth, int mis, float *sum)
{
    lro_pt_window(s, cam, W, H, 0, 0, W, H, s0, ns, seed, clamp, max_depth, mis, sum);
}
