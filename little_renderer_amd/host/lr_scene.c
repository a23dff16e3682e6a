/* lr_scene.c - seeded synthetic scenes and the camera (include/lr_scene.h). */
#include "../../include/lr_scene.h"

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t state, inc; } Pcg;

/* PCG32 (engine/base/rng.h:13-28) */
static uint32_t pcg_u32(Pcg *r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}
static void pcg_seed(Pcg *r, uint64_t seed, uint64_t stream)
{
    r->state = 0;
    r->inc = (stream << 1u) | 1u;
    pcg_u32(r);
    r->state += seed;
    pcg_u32(r);
}
static double pcg_unit(Pcg *r) { return pcg_u32(r) * (1.0 / 4294967296.0); }   /* [0, 1) */

static void random_dir(Pcg *r, double d[3])
{
    double z = 1.0 - 2.0 * pcg_unit(r), phi = 2.0 * M_PI * pcg_unit(r);
    double s = sqrt(fmax(0.0, 1.0 - z * z));
    d[0] = s * cos(phi); d[1] = s * sin(phi); d[2] = z;
}

static double lognormal(Pcg *r, double mu, double sigma)
{
    double u1 = 1.0 - pcg_unit(r), u2 = pcg_unit(r);            /* u1 in (0, 1] */
    double g = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);     /* Box-Muller */
    return exp(mu + sigma * g);
}

void lrs_gen_soup(int n, uint64_t seed, int nmat, float *tris)
{
    Pcg r;
    pcg_seed(&r, seed, 1);
    const double mu = log(0.002), sigma = 0.5;
    for (int i = 0; i < n; i++) {
        double c[3], e[2][3];
        for (int k = 0; k < 3; k++) c[k] = pcg_unit(&r);
        for (int j = 0; j < 2; j++) {
            double len = lognormal(&r, mu, sigma), d[3];
            random_dir(&r, d);
            for (int k = 0; k < 3; k++) e[j][k] = len * d[k];
        }
        float *o = tris + 12 * (size_t)i;
        for (int k = 0; k < 3; k++) {
            o[k] = (float)(c[k] - (e[0][k] + e[1][k]) / 3.0);
            o[4 + k] = (float)e[0][k];
            o[8 + k] = (float)e[1][k];
        }
        uint32_t w[3] = {0xffffffffu, 0u, (uint32_t)((nmat > 1 ? i % nmat : 0) & 63) << 26};
        for (int k = 0; k < 3; k++) { union { uint32_t u; float f; } b; b.u = w[k]; o[3 + 4 * k] = b.f; }
    }
}

static void norm3(double v[3])
{
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0) for (int k = 0; k < 3; k++) v[k] /= l;
}
static void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

void lrs_camera_make(const double from[3], const double at[3], const double up[3], double vfov_deg, double aspect,
                     float cam[12])
{
    double h = tan(vfov_deg * M_PI / 180.0 / 2.0);
    double view_h = 2.0 * h, view_w = aspect * view_h;
    double w[3] = {from[0] - at[0], from[1] - at[1], from[2] - at[2]}, u[3], v[3];
    norm3(w);
    cross3(up, w, u);
    norm3(u);
    cross3(w, u, v);
    for (int k = 0; k < 3; k++) {
        double hor = u[k] * view_w, ver = v[k] * view_h;
        cam[k] = (float)from[k];
        cam[3 + k] = (float)(from[k] - 0.5 * hor + 0.5 * ver - w[k]);
        cam[6 + k] = (float)hor;
        cam[9 + k] = (float)ver;
    }
}

/* ---- the sampler's blue-noise mask (engine/base/sampler.c:86-132): Ulichney's void-and-cluster
 * method (1993) on a 64 x 64 tile that wraps around. A pattern's "density" at a cell is the sum,
 * over the pattern's points, of a Gaussian (sigma 1.9) of the cell's toroidal offset from the
 * point. The tightest cluster is the point of largest density, the largest void the empty cell
 * of smallest density (ties: the lowest cell index). The initial pattern (a tenth of the cells,
 * drawn by PCG32 seeded (0x5eedb10e, 7)) is relaxed by moving its tightest cluster to its largest
 * void until the move is a no-op. Ranks: the initial points by removing tightest clusters from a
 * copy (the last removed gets 0), the remaining cells by filling largest voids (in order). The
 * densities are float sums taken in the order points come and go, as the reference's float
 * build takes them, so the ranks are its ranks. ---- */
enum { VC_SIDE = 64, VC_CELLS = VC_SIDE * VC_SIDE, VC_MASK = VC_SIDE - 1 };

typedef struct {
    float density[VC_CELLS];
    unsigned char set[VC_CELLS];
} VcPattern;

/* the Gaussian of each toroidal offset (ox, oy), both taken in [-32, 31] */
static float vc_gauss[VC_CELLS];

static void vc_gauss_init(void)
{
    const double sigma = (double)1.9f;   /* the float build's literal (-fsingle-precision-constant) */
    for (int oy = 0; oy < VC_SIDE; oy++) {
        const int dy = oy >= VC_SIDE / 2 ? oy - VC_SIDE : oy;
        for (int ox = 0; ox < VC_SIDE; ox++) {
            const int dx = ox >= VC_SIDE / 2 ? ox - VC_SIDE : ox;
            vc_gauss[oy * VC_SIDE + ox] = (float)exp(-(dx * dx + dy * dy) / (2 * sigma * sigma));
        }
    }
}

/* adds point p to the pattern, or takes it out: its Gaussian splatted onto every cell by offset */
static void vc_toggle(VcPattern *pat, int p)
{
    const int px = p & VC_MASK, py = p / VC_SIDE;
    const int add = !pat->set[p];
    pat->set[p] = (unsigned char)add;
    for (int oy = 0; oy < VC_SIDE; oy++) {
        float *row = pat->density + ((py + oy) & VC_MASK) * VC_SIDE;
        const float *g = vc_gauss + oy * VC_SIDE;
        for (int ox = 0; ox < VC_SIDE; ox++) {
            float *cell = row + ((px + ox) & VC_MASK);
            if (add) *cell += g[ox];
            else *cell -= g[ox];
        }
    }
}

/* the tightest cluster (of the set cells) or the largest void (of the empty ones) */
static int vc_cluster(const VcPattern *pat)
{
    int at = -1;
    float top = 0;
    for (int i = 0; i < VC_CELLS; i++)
        if (pat->set[i] && (at < 0 || pat->density[i] > top)) { at = i; top = pat->density[i]; }
    return at;
}
static int vc_void(const VcPattern *pat)
{
    int at = -1;
    float low = 0;
    for (int i = 0; i < VC_CELLS; i++)
        if (!pat->set[i] && (at < 0 || pat->density[i] < low)) { at = i; low = pat->density[i]; }
    return at;
}

void lrs_blue_noise(uint16_t *rank)
{
    static VcPattern start, work;
    vc_gauss_init();
    memset(&start, 0, sizeof start);
    Pcg r;
    pcg_seed(&r, 0x5eedb10eULL, 7);
    const int first = VC_CELLS / 10;
    int placed = 0;
    while (placed < first) {
        const int p = (int)(pcg_u32(&r) % VC_CELLS);
        if (start.set[p]) continue;
        vc_toggle(&start, p);
        placed++;
    }
    for (int moves = 0; moves < VC_CELLS; moves++) {
        const int from = vc_cluster(&start);
        vc_toggle(&start, from);
        const int to = vc_void(&start);
        vc_toggle(&start, to);
        if (to == from) break;
    }
    work = start;
    for (int k = first; k-- > 0;) {
        const int c = vc_cluster(&work);
        vc_toggle(&work, c);
        rank[c] = (uint16_t)k;
    }
    work = start;
    for (int k = first; k < VC_CELLS; k++) {
        const int v = vc_void(&work);
        vc_toggle(&work, v);
        rank[v] = (uint16_t)k;
    }
}

/* ---- the BSDF's tables (the method tools/gen_brdf_tables.py documents): E(mu, r), the
 * directional albedo of a white GGX lobe (alpha = r^2, height-correlated Smith) seen at
 * cosine mu, and B(mu, r), the same with (1 - v.h)^5 inside, on 32 x 32 nodes, each the mean
 * over 128 x 128 midpoint-stratified visible-normal samples (weight G2 / G1); E_avg, B_avg =
 * 2 * integral of E(mu) mu dmu by the trapezoid rule on 2049 points, E linearly interpolated
 * between the nodes. Stored, as the reference's table is, rounded to 6 decimals. ---- */
static double tab_g1(double c, double a) { c = fabs(c); return 2 * c / (c + sqrt(a * a + (1 - a * a) * c * c)); }
static double tab_vis(double cv, double cl, double a)
{
    const double a2 = a * a;
    const double d = cl * sqrt(cv * cv * (1 - a2) + a2) + cv * sqrt(cl * cl * (1 - a2) + a2);
    return 0.5 / d;
}
static void tab_albedo(double mu, double r, double *E, double *B)
{
    const int S = 128;
    double a = r * r;
    if (a < 1e-4) a = 1e-4;
    if (mu < 1e-4) mu = 1e-4;
    const double v[3] = {sqrt(1 - mu * mu), 0.0, mu};
    double vh[3] = {a * v[0], a * v[1], v[2]};
    const double ln = sqrt(vh[0] * vh[0] + vh[1] * vh[1] + vh[2] * vh[2]);
    for (int k = 0; k < 3; k++) vh[k] /= ln;
    double se = 0, sb = 0;
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) {
            const double u1 = (i + 0.5) / S, u2 = (j + 0.5) / S;
            const double phi = 2 * M_PI * u1;
            const double z = (1 - u2) * (1 + vh[2]) - vh[2];
            const double st = sqrt(fmax(0.0, 1 - z * z));
            const double hs[3] = {st * cos(phi) + vh[0], st * sin(phi) + vh[1], z + vh[2]};
            double h[3] = {a * hs[0], a * hs[1], fmax(hs[2], 0.0)};
            const double hl = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
            for (int k = 0; k < 3; k++) h[k] /= hl;
            const double vdh = h[0] * v[0] + h[1] * v[1] + h[2] * v[2];
            const double cl = 2 * vdh * h[2] - v[2];
            double w = 0;
            if (cl > 0) w = tab_vis(mu, fmax(cl, 1e-9), a) * 4 * mu * fmax(cl, 0.0) / tab_g1(mu, a);
            const double x = 1 - fmin(fmax(vdh, 0.0), 1.0), x5 = x * x * x * x * x;
            se += w;
            sb += w * x5;
        }
    *E = se / (S * S);
    *B = sb / (S * S);
}
static float tab_round(double x)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "%.6f", x);
    return strtof(buf, NULL);
}
static double tab_avg(const double *row)
{
    /* 2 * trapezoid of interp(fine, mu, row) * fine over fine = linspace(0, 1, 2049) */
    const int F = 2049, N = 32;
    double s = 0, prev = 0;
    for (int k = 0; k < F; k++) {
        const double x = (double)k / (F - 1);
        double pos = x * (N - 1);
        int i = (int)pos;
        if (i > N - 2) i = N - 2;
        const double t = pos - i;
        const double y = (row[i] + t * (row[i + 1] - row[i])) * x;
        if (k > 0) s += (prev + y) * 0.5 / (F - 1);
        prev = y;
    }
    return 2 * s;
}
void lrs_brdf_tables(float *out)
{
    static float cache[2112];
    static int ready;
    if (!ready) {
        double E[32][32], B[32][32];
        for (int j = 0; j < 32; j++)
            for (int i = 0; i < 32; i++) tab_albedo(i / 31.0, j / 31.0, &E[j][i], &B[j][i]);
        for (int j = 0; j < 32; j++)
            for (int i = 0; i < 32; i++) {
                cache[j * 32 + i] = tab_round(E[j][i]);
                cache[1024 + j * 32 + i] = tab_round(B[j][i]);
            }
        for (int j = 0; j < 32; j++) {
            cache[2048 + j] = tab_round(tab_avg(E[j]));
            cache[2080 + j] = tab_round(tab_avg(B[j]));
        }
        ready = 1;
    }
    for (int i = 0; i < 2112; i++) out[i] = cache[i];
}

/* ---- make_quad (scene.c:26-40), in float as the reference's default build ---- */
void lrs_make_quad(const float q[3], const float u[3], const float v[3], int material, float p[24])
{
    for (int k = 0; k < 24; k++) p[k] = 0.0f;
    const float n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    const float len = sqrtf(nn);
    float unit[3];
    for (int k = 0; k < 3; k++) unit[k] = len > 0 ? n[k] * (1.0f / len) : n[k];
    const float plane_d = unit[0] * q[0] + unit[1] * q[1] + unit[2] * q[2];
    const float inv = 1.0f / nn;
    for (int k = 0; k < 3; k++) {
        p[4 + k] = q[k];
        p[8 + k] = u[k];
        p[12 + k] = v[k];
        p[16 + k] = unit[k];
        p[20 + k] = n[k] * inv;
    }
    p[7] = 1.0f;                  /* a quad */
    p[11] = (float)material;
    p[15] = len;                  /* its area */
    p[19] = plane_d;
}

/* ---- the Cornell-style box ---- */
static void put_tri(float *o, const double a[3], const double b[3], const double c[3], int mat)
{
    for (int k = 0; k < 3; k++) {
        o[k] = (float)a[k];
        o[4 + k] = (float)(b[k] - a[k]);
        o[8 + k] = (float)(c[k] - a[k]);
    }
    uint32_t w[3] = {0xffffffffu, 0u, (uint32_t)(mat & 63) << 26};
    for (int k = 0; k < 3; k++) { union { uint32_t u; float f; } x; x.u = w[k]; o[3 + 4 * k] = x.f; }
}
/* a quad q + a u + b v as two triangles: (q, q+u, q+v) and (q+u+v, q+v, q+u), both facing u x v */
static int put_quad(float *tris, int at, const double q[3], const double u[3], const double v[3], int mat)
{
    double a[3], b[3], c[3], d[3];
    for (int k = 0; k < 3; k++) { a[k] = q[k]; b[k] = q[k] + u[k]; c[k] = q[k] + v[k]; d[k] = q[k] + u[k] + v[k]; }
    put_tri(tris + 12 * (size_t)at, a, b, c, mat);
    put_tri(tris + 12 * (size_t)(at + 1), d, c, b, mat);
    return at + 2;
}
static void rot_y(const double p[3], double deg, double o[3])
{
    const double a = deg * M_PI / 180.0, c = cos(a), s = sin(a);
    o[0] = c * p[0] + s * p[2];
    o[1] = p[1];
    o[2] = -s * p[0] + c * p[2];
}
/* add_box (scene.c:379-389): 6 faces of a box (0,0,0)..size turned by deg about y, moved by off */
static int put_box(float *tris, int at, const double size[3], double deg, const double off[3], int *mat, double sc)
{
    const double dx[3] = {size[0], 0, 0}, dy[3] = {0, size[1], 0}, dz[3] = {0, 0, size[2]};
    const double corner[6][3] = {{0, 0, 0}, {0, 0, size[2]}, {0, 0, 0}, {size[0], 0, 0}, {0, 0, 0}, {0, size[1], 0}};
    const double *eu[6] = {dx, dx, dz, dz, dx, dx}, *ev[6] = {dy, dy, dy, dy, dz, dz};
    for (int i = 0; i < 6; i++) {
        double q[3], u[3], v[3];
        rot_y(corner[i], deg, q);
        rot_y(eu[i], deg, u);
        rot_y(ev[i], deg, v);
        for (int k = 0; k < 3; k++) { q[k] = (q[k] + off[k]) * sc; u[k] *= sc; v[k] *= sc; }
        at = put_quad(tris, at, q, u, v, (*mat)++);
    }
    return at;
}

void lrs_cornell(uint64_t seed, int w, int h, float *tris, float *prims, LrhMaterial *mats, float cam[12])
{
    const double sc = 1.0 / 555.0;
    Pcg r;
    pcg_seed(&r, seed, 1);
    for (int i = 0; i < 17; i++) {
        LrhMaterial *m = &mats[i];
        memset(m, 0, sizeof(*m));
        m->type = LRH_MAT_DIFFUSE;
        for (int k = 0; k < 3; k++) m->albedo[k] = (float)(0.2 + 0.6 * pcg_unit(&r));
        m->ior = 1.5f;
        m->roughness = 1.0f;
    }
    LrhMaterial *lm = &mats[17];
    memset(lm, 0, sizeof(*lm));
    lm->type = LRH_MAT_LIGHT;
    lm->emission[0] = lm->emission[1] = lm->emission[2] = 15.0f;
    lm->ior = 1.0f;
    lm->roughness = 1.0f;
    /* the walls (scene.c:449-453): left, right, floor, ceiling, back */
    const double wq[5][3] = {{555, 0, 0}, {0, 0, 0}, {0, 0, 0}, {555, 555, 555}, {0, 0, 555}};
    const double wu[5][3] = {{0, 555, 0}, {0, 555, 0}, {555, 0, 0}, {-555, 0, 0}, {555, 0, 0}};
    const double wv[5][3] = {{0, 0, 555}, {0, 0, 555}, {0, 0, 555}, {0, 0, -555}, {0, 555, 0}};
    int at = 0;
    for (int i = 0; i < 5; i++) {
        double q[3], u[3], v[3];
        for (int k = 0; k < 3; k++) { q[k] = wq[i][k] * sc; u[k] = wu[i][k] * sc; v[k] = wv[i][k] * sc; }
        at = put_quad(tris, at, q, u, v, i);
    }
    int mat = 5;
    const double tall[3] = {165, 330, 165}, tall_off[3] = {265, 0, 295};
    const double shrt[3] = {165, 165, 165}, shrt_off[3] = {130, 0, 65};
    at = put_box(tris, at, tall, 15.0, tall_off, &mat, sc);
    at = put_box(tris, at, shrt, -18.0, shrt_off, &mat, sc);
    /* the lamp: 0.25 x 0.25, centred under the ceiling, u x v pointing down */
    const float lq[3] = {0.625f, (float)(554.0 * sc), 0.625f}, lu[3] = {-0.25f, 0.0f, 0.0f}, lv[3] = {0.0f, 0.0f, -0.25f};
    lrs_make_quad(lq, lu, lv, 17, prims);
    const double from[3] = {278 * sc, 278 * sc, -800 * sc}, to[3] = {278 * sc, 278 * sc, 0}, up[3] = {0, 1, 0};
    lrs_camera_make(from, to, up, 40.0, (double)w / (double)h, cam);
}

/* ---- the batched and serving workloads ---- */
void lrs_batch_scene(int n, uint64_t seed, float *tris, float *prims, LrhMaterial *mats)
{
    Pcg r;
    pcg_seed(&r, seed, 3);
    for (int i = 0; i < LRS_BATCH_SOUP_MATS; i++) {
        LrhMaterial *m = &mats[i];
        memset(m, 0, sizeof(*m));
        for (int k = 0; k < 3; k++) m->albedo[k] = (float)(0.2 + 0.7 * pcg_unit(&r));
        m->roughness = (float)(0.05 + 0.95 * pcg_unit(&r));
        m->metallic = pcg_unit(&r) < 0.4 ? 1.0f : 0.0f;
        m->ior = 1.5f;
        m->type = LRH_MAT_PBR;
        if (i % 7 == 6) {           /* two of the fourteen: clear glass */
            m->type = LRH_MAT_GLASS;
            m->albedo[0] = m->albedo[1] = m->albedo[2] = 0.95f;
            m->metallic = 0.0f;
            m->roughness = 0.0f;
            m->transmission = 1.0f;
        }
    }
    LrhMaterial *fl = &mats[LRS_BATCH_SOUP_MATS], *lm = &mats[LRS_BATCH_SOUP_MATS + 1];
    memset(fl, 0, sizeof(*fl));
    fl->type = LRH_MAT_DIFFUSE;
    fl->albedo[0] = fl->albedo[1] = fl->albedo[2] = 0.5f;
    fl->ior = 1.5f;
    fl->roughness = 1.0f;
    memset(lm, 0, sizeof(*lm));
    lm->type = LRH_MAT_LIGHT;
    lm->emission[0] = lm->emission[1] = lm->emission[2] = 15.0f;
    lm->ior = 1.0f;
    lm->roughness = 1.0f;
    lrs_gen_soup(n, seed, LRS_BATCH_SOUP_MATS, tris);
    const double fq[3] = {-2.0, -0.05, -2.0}, fu[3] = {0.0, 0.0, 5.0}, fv[3] = {5.0, 0.0, 0.0};
    put_quad(tris, n, fq, fu, fv, LRS_BATCH_SOUP_MATS);
    const float lq[3] = {0.75f, 1.6f, 0.75f}, lu[3] = {-0.5f, 0.0f, 0.0f}, lv[3] = {0.0f, 0.0f, -0.5f};
    lrs_make_quad(lq, lu, lv, LRS_BATCH_SOUP_MATS + 1, prims);
}

void lrs_batch_camera(uint64_t seed, int w, int h, float cam[12])
{
    Pcg r;
    pcg_seed(&r, seed, 4);
    double d[3];
    do random_dir(&r, d); while (d[1] < 0.15);
    const double at[3] = {0.5, 0.5, 0.5}, up[3] = {0, 1, 0};
    const double from[3] = {0.5 + 2.2 * d[0], 0.5 + 2.2 * d[1], 0.5 + 2.2 * d[2]};
    lrs_camera_make(from, at, up, 40.0, (double)w / (double)h, cam);
}

void lrs_batch_plan(int njobs, uint64_t seed, int spp_lo, int *ntris, int *w, int *h, int *spp)
{
    static const int res[4][2] = {{256, 256}, {512, 512}, {1280, 720}, {1920, 1080}};
    static const int spps[4] = {1, 4, 16, 64};
    Pcg r;
    pcg_seed(&r, seed, 5);
    const int first = spp_lo == 4 ? 1 : 0;
    for (int i = 0; i < njobs; i++) {
        ntris[i] = (int)floor(exp(log(1000.0) + (log(500000.0) - log(1000.0)) * pcg_unit(&r)) + 0.5);
        const int k = (int)(pcg_u32(&r) % 4u);
        w[i] = res[k][0];
        h[i] = res[k][1];
        spp[i] = spps[first + (int)(pcg_u32(&r) % (uint32_t)(4 - first))];
    }
}

/* ---- the sky (engine/scene/env.c) ---- */
static const float SKY_PI = 3.14159265358979323846f;

/* map_dir (env.c:48-52): the direction of map point (u, v), y up, u = 0.5 looking along +z */
static void sky_dir(float u, float v, float d[3])
{
    const float ph = (u - 0.5f) * (2.0f * SKY_PI), th = v * SKY_PI;
    const float st = sinf(th);
    d[0] = st * sinf(ph);
    d[1] = cosf(th);
    d[2] = st * cosf(ph);
}

void lrs_sky(int w, float sun_elev, float sun_azim, const float sun_le[3], float *rgb)
{
    const int h = w / 2;
    const float el = sun_elev * SKY_PI / 180.0f, az = sun_azim * SKY_PI / 180.0f;
    const float sun[3] = {cosf(el) * sinf(az), sinf(el), cosf(el) * cosf(az)};
    const float sun_cos = cosf(1.0f * SKY_PI / 180.0f);
    const float zen[3] = {0.18f, 0.32f, 0.75f}, hor[3] = {0.75f, 0.8f, 0.85f}, glow[3] = {1.0f, 0.9f, 0.7f};
    const float gnd_a[3] = {0.7f, 0.72f, 0.75f}, gnd_b[3] = {0.12f, 0.11f, 0.1f};
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float d[3], c[3];
            sky_dir((x + 0.5f) / w, (y + 0.5f) / h, d);
            if (d[1] >= 0) {
                const float t = 1 - d[1], t4 = t * t * t * t;
                const float cs = d[0] * sun[0] + d[1] * sun[1] + d[2] * sun[2];
                const float g = 0.6f * powf(cs > 0 ? cs : 0, 32);
                for (int k = 0; k < 3; k++) {
                    c[k] = zen[k] * (1 - t4) + hor[k] * t4;
                    c[k] = c[k] + glow[k] * g;
                    if (cs > sun_cos) c[k] = c[k] + sun_le[k];
                }
            } else {
                float t = -d[1] * 4;
                t = t < 1 ? t : 1;
                for (int k = 0; k < 3; k++) c[k] = gnd_a[k] * (1 - t) + gnd_b[k] * t;
            }
            float *o = rgb + 3 * ((size_t)y * w + x);
            o[0] = c[0];
            o[1] = c[1];
            o[2] = c[2];
        }
}

int lrs_sky_tables(const float *rgb, int w, int h, float *marg, float *cond, float *prob, double *sum_out)
{
    if (!rgb || w < 2 || h < 1 || !marg || !cond || !prob) return -1;
    float *lum = malloc(sizeof(float) * (size_t)w * h), *wm = malloc(sizeof(float) * (size_t)w * h);
    float *rs = malloc(sizeof(float) * (size_t)h);
    double *rows = malloc(sizeof(double) * (size_t)h);
    if (!lum || !wm || !rs || !rows) {
        free(lum); free(wm); free(rs); free(rows);
        return -1;
    }
    /* The sampling weight of a pixel (env.c:125-162, pixel_weight): the largest Rec. 709
       luminance over its 3 x 3 neighbourhood, the map wrapping around in x and stopping at the
       poles in y, never below 0; each row's weights are taken times sin(theta) at its centre.
       Luminances first, then each pixel's nine neighbours in one pass. */
    for (size_t i = 0; i < (size_t)w * h; i++)
        lum[i] = 0.2126f * rgb[3 * i] + 0.7152f * rgb[3 * i + 1] + 0.0722f * rgb[3 * i + 2];
    for (int y = 0; y < h; y++) {
        const int rows3[3] = {y > 0 ? y - 1 : y, y, y + 1 < h ? y + 1 : y};
        for (int x = 0; x < w; x++) {
            const int cols3[3] = {(x + w - 1) % w, x, (x + 1) % w};
            float m = 0;
            for (int j = 0; j < 3; j++)
                for (int i = 0; i < 3; i++) m = fmaxf(m, lum[(size_t)rows3[j] * w + cols3[i]]);
            wm[(size_t)y * w + x] = m;
        }
        rs[y] = sinf((y + 0.5f) * SKY_PI / h);
    }
    /* prepare_base's sampling tables (env.c:343-364): each row's cumulative distribution over
       its pixels, the rows' over the map, weights lum_max * sin(theta) summed in double */
    double sum = 0;
    for (int y = 0; y < h; y++) {
        float *c = cond + (size_t)y * (w + 1);
        double acc = 0;
        c[0] = 0;
        for (int x = 0; x < w; x++) {
            acc += (double)(wm[(size_t)y * w + x] * rs[y]);
            c[x + 1] = (float)acc;
        }
        rows[y] = acc;
        for (int x = 1; x <= w; x++) c[x] = acc > 0 ? (float)(c[x] / acc) : (float)x / w;
        c[w] = 1;
        sum += acc;
    }
    double acc = 0;
    marg[0] = 0;
    for (int y = 0; y < h; y++) {
        acc += rows[y];
        marg[y + 1] = sum > 0 ? (float)(acc / sum) : (float)(y + 1) / h;
    }
    marg[h] = 1;
    /* pixel_prob (env.c:164-167), per pixel */
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            prob[(size_t)y * w + x] = sum > 0 ? (float)((double)(wm[(size_t)y * w + x] * rs[y]) / sum) : 0.0f;
    if (sum_out) *sum_out = sum;
    free(lum); free(wm); free(rs); free(rows);
    return 0;
}

int lrs_affine_inverse(const float fwd[12], float inv[12])
{
    /* [A | t]^-1 = [A^-1 | -A^-1 t], in double, by the adjugate */
    double a[3][3], t[3];
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) a[r][c] = fwd[4 * r + c];
        t[r] = fwd[4 * r + 3];
    }
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (!(fabs(det) > 1e-30)) return -1;
    double b[3][3];
    b[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    b[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    b[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    b[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    b[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    b[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    b[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    b[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    b[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) inv[4 * r + c] = (float)b[r][c];
        inv[4 * r + 3] = (float)(-(b[r][0] * t[0] + b[r][1] * t[1] + b[r][2] * t[2]));
    }
    return 0;
}

/* ---- meshes of the instanced workload (object space, material in the top bits) ---- */
static void put_tri_f(float *o, const double a[3], const double b[3], const double c[3], int mat) { put_tri(o, a, b, c, mat); }

int lrs_uv_sphere(int slices, int stacks, int mat, float *tris)
{
    int at = 0;
    for (int j = 0; j < stacks; j++) {
        const double t0 = M_PI * j / stacks, t1 = M_PI * (j + 1) / stacks;
        for (int i = 0; i < slices; i++) {
            const double p0 = 2 * M_PI * i / slices, p1 = 2 * M_PI * (i + 1) / slices;
            const double a[3] = {sin(t0) * cos(p0), cos(t0), sin(t0) * sin(p0)}, b[3] = {sin(t0) * cos(p1), cos(t0), sin(t0) * sin(p1)};
            const double c[3] = {sin(t1) * cos(p0), cos(t1), sin(t1) * sin(p0)}, d[3] = {sin(t1) * cos(p1), cos(t1), sin(t1) * sin(p1)};
            /* outward (u x v away from the centre) */
            if (j > 0) put_tri_f(tris + 12 * (size_t)at++, a, b, c, mat);
            if (j < stacks - 1) put_tri_f(tris + 12 * (size_t)at++, b, d, c, mat);
        }
    }
    return at;
}

int lrs_box_mesh(int mat, float *tris)
{
    int at = 0;
    for (int ax = 0; ax < 3; ax++)
        for (int s = -1; s <= 1; s += 2) {
            const int u = (ax + 1) % 3, v = (ax + 2) % 3;
            double q[3] = {0, 0, 0}, du[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
            q[ax] = s;
            q[u] = -1;
            q[v] = -1;
            du[u] = 2;
            dv[v] = 2;
            if (s < 0) {   /* keep u x v pointing out */
                double t[3];
                memcpy(t, du, sizeof(t));
                memcpy(du, dv, sizeof(t));
                memcpy(dv, t, sizeof(t));
            }
            at = put_quad(tris, at, q, du, dv, mat);
        }
    return at;
}

int lrs_hull(int npts, uint64_t seed, int mat, float *tris)
{
    /* the convex hull of npts points drawn uniformly on the unit sphere: with every point on
       the hull, a face is a triple with all the other points on one side (O(n^4), n small) */
    Pcg r;
    pcg_seed(&r, seed, 6);
    double (*p)[3] = malloc(sizeof(double) * 3 * (size_t)npts);
    if (!p) return -1;
    for (int i = 0; i < npts; i++) random_dir(&r, p[i]);
    int at = 0;
    for (int i = 0; i < npts; i++)
        for (int j = i + 1; j < npts; j++)
            for (int k = j + 1; k < npts; k++) {
                double e1[3], e2[3], n[3];
                for (int q = 0; q < 3; q++) { e1[q] = p[j][q] - p[i][q]; e2[q] = p[k][q] - p[i][q]; }
                cross3(e1, e2, n);
                const double d0 = n[0] * p[i][0] + n[1] * p[i][1] + n[2] * p[i][2];
                int pos = 0, neg = 0;
                for (int m = 0; m < npts && !(pos && neg); m++) {
                    if (m == i || m == j || m == k) continue;
                    const double s = n[0] * p[m][0] + n[1] * p[m][1] + n[2] * p[m][2] - d0;
                    if (s > 0) pos = 1;
                    else neg = 1;
                }
                if (pos && neg) continue;
                if (pos) put_tri_f(tris + 12 * (size_t)at++, p[i], p[k], p[j], mat);   /* the others above: flip */
                else put_tri_f(tris + 12 * (size_t)at++, p[i], p[j], p[k], mat);
            }
    free(p);
    return at;
}

int lrs_blob(int ntris, uint64_t seed, int mat, float *tris)
{
    /* a UV sphere of about ntris triangles, its radius bumped by a few seeded waves */
    int stacks = (int)floor(sqrt(ntris / 4.0) + 0.5) + 1;
    if (stacks < 3) stacks = 3;
    int slices = (int)floor(ntris / (2.0 * (stacks - 1)) + 0.5);
    if (slices < 3) slices = 3;
    const int n = lrs_uv_sphere(slices, stacks, mat, tris);
    Pcg r;
    pcg_seed(&r, seed, 7);
    double k[4][3], ph[4], amp[4];
    for (int w = 0; w < 4; w++) {
        random_dir(&r, k[w]);
        const double f = 1.0 + 3.0 * pcg_unit(&r);
        for (int q = 0; q < 3; q++) k[w][q] *= f;
        ph[w] = 2 * M_PI * pcg_unit(&r);
        amp[w] = 0.05 + 0.1 * pcg_unit(&r);
    }
    for (int t = 0; t < n; t++) {
        float *o = tris + 12 * (size_t)t;
        double v[3][3];
        for (int q = 0; q < 3; q++) { v[0][q] = o[q]; v[1][q] = o[q] + o[4 + q]; v[2][q] = o[q] + o[8 + q]; }
        for (int c = 0; c < 3; c++) {
            double s = 1.0;
            for (int w = 0; w < 4; w++) s += amp[w] * sin(k[w][0] * v[c][0] + k[w][1] * v[c][1] + k[w][2] * v[c][2] + ph[w]);
            for (int q = 0; q < 3; q++) v[c][q] *= s;
        }
        uint32_t wd;
        memcpy(&wd, &o[11], 4);
        put_tri_f(o, v[0], v[1], v[2], (int)(wd >> 26));
    }
    return n;
}
