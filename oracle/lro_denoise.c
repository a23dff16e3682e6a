/* lro_denoise.c - TEST INFRASTRUCTURE ONLY: the oracle's denoiser, a plain-C restatement of the
 * reference's edge-avoiding a-trous filter (engine/gpu/denoise_gpu.c:1-20, 73-99;
 * engine/gpu/shaders/trace/denoise.wgsl:1-205) on the inputs lrh_denoise takes: per pixel the
 * sums of the samples (float4 rgb, count), the moments of their irradiance luminance (float2),
 * and the guide (8 floats: normal, distance | albedo, kind). Its arithmetic is written out in the
 * order the device's is, in float, with the same packing between passes: halves rounded to
 * nearest even, the normal on an octahedron in two 16-bit snorms, the kind in the distance's
 * sign. Its exp and sqrt are glibc's, so a weight may differ in the last bit from the device's,
 * and a packed half, rarely, by one step.
 *
 * Pinned by tests/test_denoise_oracle.py: the packing against the IEEE-754 half routine numpy
 * carries, the octahedral map against a closed round trip, a constant image passing unchanged,
 * a one-pixel impulse spreading by the B3 spline's weights, edges (normal, distance, kind)
 * stopping it, and the variance falling as the passes go. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lro_pt.h"

typedef struct { float x, y, z; } F3;
static inline F3 f3(float x, float y, float z) { F3 r = {x, y, z}; return r; }
static inline float fmin_(float a, float b) { return b != b ? a : (a < b ? a : b); }
static inline float fmax_(float a, float b) { return b != b ? a : (a > b ? a : b); }
static inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float luma(F3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
static inline F3 norm3(F3 a)
{
    float l = sqrtf(dot3(a, a));
    if (!(l > 0)) return a;
    float k = 1.0f / l;
    return f3(a.x * k, a.y * k, a.z * k);
}

/* ---- IEEE-754 binary16, round to nearest even ---- */
uint16_t lro_f2h(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);          /* 65520 and above: infinity */
    if (ax < 0x38800000u) {                                            /* below 2^-14: a subnormal */
        if (ax < 0x33000000u) return (uint16_t)sign;                   /* 2^-25 and below: zero */
        uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
        int shift = 126 - (int)e;
        uint32_t mh = m >> shift, rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (mh & 1u))) mh++;
        return (uint16_t)(sign | mh);
    }
    uint32_t h = (ax >> 13) - (112u << 10), rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}
float lro_h2f(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
    if (e == 0) {
        if (m == 0) x = sign;
        else {   /* a subnormal: m * 2^-24 */
            float f = (float)m * (1.0f / 16777216.0f);
            memcpy(&x, &f, 4);
            x |= sign;
        }
    } else if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else x = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* irradiance and variance as four halves (denoise.wgsl:12-17) */
typedef struct { uint16_t h[4]; } IV;
static IV put_iv(F3 c, float v)
{
    IV p;
    p.h[0] = lro_f2h(fmin_(c.x, 6e4f));
    p.h[1] = lro_f2h(fmin_(c.y, 6e4f));
    p.h[2] = lro_f2h(fmin_(c.z, 6e4f));
    p.h[3] = lro_f2h(fmin_(fmax_(v, -1.0f), 6e4f));
    return p;
}
static void get_iv(IV p, float o[4])
{
    for (int k = 0; k < 4; k++) o[k] = lro_h2f(p.h[k]);
}

static uint16_t snorm16(float x)
{
    float c = fmin_(fmax_(x, -1.0f), 1.0f);
    return (uint16_t)(int16_t)floorf(0.5f + 32767.0f * c);
}
static float unsnorm16(uint16_t u) { return fmax_((float)(int16_t)u / 32767.0f, -1.0f); }

/* the normal on an octahedron (denoise.wgsl:21-33) */
void lro_oct_pack(const float n[3], uint16_t out[2])
{
    float s = fmax_(fabsf(n[0]) + fabsf(n[1]) + fabsf(n[2]), 1e-6f);
    float px = n[0] / s, py = n[1] / s;
    if (n[2] < 0.0f) {
        float fx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
        float fy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
        px = fx;
        py = fy;
    }
    out[0] = snorm16(px);
    out[1] = snorm16(py);
}
void lro_oct_unpack(const uint16_t in[2], float n[3])
{
    F3 v = f3(unsnorm16(in[0]), unsnorm16(in[1]), 0.0f);
    v.z = 1.0f - fabsf(v.x) - fabsf(v.y);
    float t = fmax_(-v.z, 0.0f);
    v.x += v.x >= 0.0f ? -t : t;
    v.y += v.y >= 0.0f ? -t : t;
    v = norm3(v);
    n[0] = v.x;
    n[1] = v.y;
    n[2] = v.z;
}

typedef struct { uint16_t oct[2]; float e; } GP;   /* the packed guide */
typedef struct { F3 n; float z; int kind; } Gd;
static GP put_g(const float *g)
{
    GP p;
    lro_oct_pack(g, p.oct);
    p.e = g[7] == 0.0f ? g[3] : -2e38f;
    return p;
}
static Gd get_g(GP p)
{
    Gd g;
    float n[3];
    g.kind = p.e >= 0.0f ? 0 : 1;
    g.z = p.e >= 0.0f ? p.e : -p.e - 1.0f;
    lro_oct_unpack(p.oct, n);
    g.n = f3(n[0], n[1], n[2]);
    return g;
}

static F3 irradiance(const float *a, const float *alb)
{
    float k = fmax_(a[3], 1.0f);
    F3 mean = f3(a[0] / k, a[1] / k, a[2] / k);
    return alb[3] == 0.0f ? f3(mean.x / fmax_(alb[0], 1e-3f), mean.y / fmax_(alb[1], 1e-3f), mean.z / fmax_(alb[2], 1e-3f))
                          : mean;
}

static uint32_t hash32(uint32_t x)   /* util.wgsl:6-10 (Wellons' lowbias32) */
{
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

static void prep(const float *accum, const float *mom, const float *guide, int W, int H, GP *gd, IV *dst)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t k = (size_t)y * W + x;
            const float *a = accum + 4 * k, *gn = guide + 8 * k, *alb = guide + 8 * k + 4;
            gd[k] = put_g(gn);
            if (a[3] == 0.0f) {
                dst[k] = put_iv(f3(0, 0, 0), -1.0f);
                continue;
            }
            F3 irr = irradiance(a, alb);
            if (alb[3] != 0.0f) {
                dst[k] = put_iv(irr, 0.0f);
                continue;
            }
            float v = 0.0f;
            if (a[3] >= 4.0f) {
                float el = mom[2 * k] / a[3];
                v = fmax_(mom[2 * k + 1] / a[3] - el * el, 0.0f) / a[3];
            } else {
                float s1 = 0, s2 = 0, c = 0;
                for (int dy = -3; dy <= 3; dy++)
                    for (int dx = -3; dx <= 3; dx++) {
                        int qx = x + dx, qy = y + dy;
                        if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                        size_t q = (size_t)qy * W + qx;
                        const float *aq = accum + 4 * q, *albq = guide + 8 * q + 4, *gq = guide + 8 * q;
                        if (aq[3] == 0.0f || albq[3] != alb[3]) continue;
                        if (dot3(f3(gq[0], gq[1], gq[2]), f3(gn[0], gn[1], gn[2])) < 0.9f) continue;
                        float l = luma(irradiance(aq, albq));
                        s1 += l;
                        s2 += l * l;
                        c += 1.0f;
                    }
                float el = s1 / c;
                v = fmax_(s2 / c - el * el, 0.0f);
            }
            dst[k] = put_iv(irr, v);
        }
}

static void atrous(const IV *src, const GP *gd, const float *guide, int W, int H, int step, int last, IV *dst, float *out)
{
    static const float hb[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t k = (size_t)y * W + x;
            float c[4];
            get_iv(src[k], c);
            Gd gp = get_g(gd[k]);
            if (c[3] < 0.0f || gp.kind == 1) {
                if (last) {
                    float *o = out + 4 * k;
                    if (c[3] < 0.0f) o[0] = o[1] = o[2] = o[3] = 0.0f;
                    else { o[0] = c[0]; o[1] = c[1]; o[2] = c[2]; o[3] = 1.0f; }
                } else dst[k] = src[k];
                continue;
            }
            float v = 0.0f, vw = 0.0f;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int qx = x + dx < 0 ? 0 : (x + dx > W - 1 ? W - 1 : x + dx);
                    int qy = y + dy < 0 ? 0 : (y + dy > H - 1 ? H - 1 : y + dy);
                    float s[4];
                    get_iv(src[(size_t)qy * W + qx], s);
                    if (s[3] < 0.0f) continue;
                    float w = (dx != 0 ? 0.25f : 0.5f) * (dy != 0 ? 0.25f : 0.5f);
                    v += w * s[3];
                    vw += w;
                }
            v /= fmax_(vw, 1e-6f);
            float lp = luma(f3(c[0], c[1], c[2]));
            float zl = get_g(gd[(size_t)y * W + (x - 1 < 0 ? 0 : x - 1)]).z;
            float zr = get_g(gd[(size_t)y * W + (x + 1 > W - 1 ? W - 1 : x + 1)]).z;
            float zd = get_g(gd[(size_t)(y - 1 < 0 ? 0 : y - 1) * W + x]).z;
            float zu = get_g(gd[(size_t)(y + 1 > H - 1 ? H - 1 : y + 1) * W + x]).z;
            float gx = fmin_(fabsf(gp.z - zl), fabsf(gp.z - zr)), gy = fmin_(fabsf(gp.z - zd), fabsf(gp.z - zu));
            uint32_t seed = hash32(((uint32_t)k * 0x9e3779b9u) ^ ((uint32_t)step * 0x85ebca6bu));
            int jit = step / 2;
            uint32_t jn = (uint32_t)(2 * jit + 1);
            F3 sum = f3(0, 0, 0);
            float vsum = 0.0f, wsum = 0.0f;
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int ox = (i - 2) * step, oy = (j - 2) * step;
                    if (jit > 0 && (i != 2 || j != 2)) {
                        seed = hash32(seed);
                        ox += (int)(((seed & 0xffffu) * jn) >> 16) - jit;
                        oy += (int)(((seed >> 16) * jn) >> 16) - jit;
                    }
                    int qx = x + ox, qy = y + oy;
                    if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                    size_t q = (size_t)qy * W + qx;
                    float s[4];
                    get_iv(src[q], s);
                    float w = hb[i] * hb[j];
                    if (q != k) {
                        Gd gq = get_g(gd[q]);
                        if (s[3] < 0.0f || gq.kind != gp.kind) continue;
                        w *= expf(-fabsf(lp - luma(f3(s[0], s[1], s[2]))) / (4.0f * sqrtf(fmax_(v, s[3])) + 1e-6f));
                        float wn = fmax_(dot3(gp.n, gq.n), 0.0f);
                        for (int e = 0; e < 7; e++) wn *= wn;
                        w *= wn;
                        float expect = gx * fabsf((float)ox) + gy * fabsf((float)oy);
                        w *= expf(-fabsf(gp.z - gq.z) / (1.0f * expect + 1e-3f * gp.z + 1e-6f));
                    }
                    sum = f3(sum.x + s[0] * w, sum.y + s[1] * w, sum.z + s[2] * w);
                    vsum += w * w * s[3];
                    wsum += w;
                }
            float inv = 1.0f / wsum;
            F3 o = f3(sum.x * inv, sum.y * inv, sum.z * inv);
            if (last) {
                const float *alb = guide + 8 * k + 4;
                float *op = out + 4 * k;
                op[0] = o.x * alb[0];
                op[1] = o.y * alb[1];
                op[2] = o.z * alb[2];
                op[3] = 1.0f;
            } else dst[k] = put_iv(o, vsum / (wsum * wsum));
        }
}

void lro_denoise(const float *accum, const float *moments, const float *guide, int W, int H, int passes, float *out)
{
    size_t n = (size_t)W * H;
    IV *a = malloc(n * sizeof(IV)), *b = malloc(n * sizeof(IV));
    GP *gd = malloc(n * sizeof(GP));
    if (!a || !b || !gd) {
        free(a);
        free(b);
        free(gd);
        return;
    }
    prep(accum, moments, guide, W, H, gd, a);
    for (int i = 1; i <= passes; i++) {
        const IV *src = i % 2 ? a : b;
        IV *dst = i % 2 ? b : a;
        atrous(src, gd, guide, W, H, 1 << (i - 1), i == passes, dst, out);
    }
    free(a);
    free(b);
    free(gd);
}
