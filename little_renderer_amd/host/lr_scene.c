/* lr_scene.c - seeded synthetic scenes and the camera (include/lr_scene.h). */
#include "../../include/lr_scene.h"

#include <math.h>
#include <stddef.h>

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
