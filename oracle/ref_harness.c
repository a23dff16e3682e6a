/* ref_harness.c - plain-array entry points into the reference's own CPU engine, compiled
 * together with its unmodified sources (engine/base, formats, scene, rt) by the "ref" target
 * of oracle/Makefile into oracle/_ref/. TEST INFRASTRUCTURE ONLY: it lets the tests check the
 * oracle's restatement against the reference itself, and the benchmark time the reference's
 * CPU path as its baseline. It calls the reference's functions; it replaces none of them. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "base/rng.h"
#include "base/sampler.h"
#include "rt/bvh.h"
#include "scene/scene.h"

/* PCG32 outputs (rng.h:13-28) */
void refh_pcg32(uint64_t seed, uint64_t stream, int n, uint32_t *out)
{
    Rng r;
    rng_seed(&r, seed, stream);
    for (int i = 0; i < n; i++) out[i] = rng_u32(&r);
}

/* rng_real (rng.h:31-37) as this build's real, widened to double */
void refh_rng_real(uint64_t seed, uint64_t stream, int n, double *out)
{
    Rng r;
    rng_seed(&r, seed, stream);
    for (int i = 0; i < n; i++) out[i] = (double)rng_real(&r);
}

/* sampler_2d (sampler.c:143-160) for a list of (x, y, index, dim), scramble `seed` */
void refh_sampler_2d(int n, const int *x, const int *y, const uint32_t *index, const int *dim, uint32_t seed,
                     double *out)
{
    for (int i = 0; i < n; i++) {
        Rng rng;
        rng_seed(&rng, 1, 1);
        Sampler s = sampler_make(&rng, x[i], y[i], index[i], seed, 1);
        real a, b;
        sampler_2d(&s, dim[i], &a, &b);
        out[2 * i] = (double)a;
        out[2 * i + 1] = (double)b;
    }
}

static void tris_in(Scene *s, const float *tris, int n)
{
    memset(s, 0, sizeof(*s));
    s->tris = calloc((size_t)(n > 0 ? n : 1), sizeof(Triangle));
    for (int i = 0; i < n; i++) {
        const float *t = tris + 12 * (size_t)i;
        Triangle *q = &s->tris[i];
        q->q = vec3(t[0], t[1], t[2]);
        q->u = vec3(t[4], t[5], t[6]);
        q->v = vec3(t[8], t[9], t[10]);
        for (int k = 0; k < 3; k++) { q->n[k] = -1; q->t[k] = -1; }
        q->material = 0;
    }
    s->ntris = n;
    s->light = -1;
}

/* bvh_build (bvh.c:447, the scene's shapes = these triangles, their triangle_bounds) with
 * max_leaf, bins, SAH; its nodes in the GPU layout (put_nodes, trace_gpu.c:883-895) and refs.
 * Returns the node count; depth and cost (bvh_cost) out. */
int refh_bvh_tris(const float *tris, int n, int max_leaf, int bins, uint32_t *nodes, int *refs, int *depth, double *cost)
{
    Scene s;
    tris_in(&s, tris, n);
    Bvh b;
    bvh_build(&b, &s, max_leaf, bins, BVH_SPLIT_SAH);
    for (int i = 0; i < b.nnodes; i++) {
        const BvhNode *nd = &b.nodes[i];
        float lo[3] = {(float)nd->box.lo.x, (float)nd->box.lo.y, (float)nd->box.lo.z};
        float hi[3] = {(float)nd->box.hi.x, (float)nd->box.hi.y, (float)nd->box.hi.z};
        uint32_t *o = nodes + 8 * (size_t)i;
        memcpy(o, lo, 12);
        memcpy(o + 4, hi, 12);
        o[3] = nd->count > 0 ? (uint32_t)nd->start : (uint32_t)nd->left;
        o[7] = nd->count > 0 ? (0x80000000u | (uint32_t)nd->count) : (uint32_t)nd->right;
    }
    for (int i = 0; i < b.nrefs; i++) refs[i] = b.refs[i];
    *depth = b.depth;
    *cost = (double)bvh_cost(&b);
    int nn = b.nnodes;
    bvh_free(&b);
    free(s.tris);
    return nn;
}

/* the reference CPU's scene of these triangles with its SAH BVH (bvh_build, max leaf 4, 16
 * bins: accel_default_params, accel.c:28-31), 4-wide when wide (bvh_make_wide) */
typedef struct { Scene s; Bvh b; } RefTris;

void *refh_tris_new(const float *tris, int n, int wide)
{
    RefTris *r = calloc(1, sizeof(RefTris));
    tris_in(&r->s, tris, n);
    bvh_build(&r->b, &r->s, 4, 16, BVH_SPLIT_SAH);
    if (wide) bvh_make_wide(&r->b);
    return r;
}

void refh_tris_free(void *h)
{
    RefTris *r = h;
    if (!r) return;
    bvh_free(&r->b);
    free(r->s.tris);
    free(r);
}

/* bvh_intersect (bvh.c:608-663; wide: wide_intersect :543-604) of rays (o.xyz tmin d.xyz
 * tmax), OpenMP over the rays: hits (t, id, b1, b2), id 0xffffffff on a miss */
void refh_tris_trace(void *h, const float *rays, int nrays, int any, float *hits)
{
    RefTris *R = h;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < nrays; i++) {
        const float *r = rays + 8 * (size_t)i;
        Ray ray = {vec3(r[0], r[1], r[2]), vec3(r[4], r[5], r[6])};
        Hit hh;
        float *o = hits + 4 * (size_t)i;
        if (bvh_intersect(&R->b, &R->s, &ray, r[3], r[7], &hh, any)) {
            o[0] = (float)hh.t;
            uint32_t id = (uint32_t)hh.prim;
            memcpy(&o[1], &id, 4);
            o[2] = (float)hh.uv.x;
            o[3] = (float)hh.uv.y;
        } else {
            uint32_t none = 0xffffffffu;
            o[0] = r[7];
            memcpy(&o[1], &none, 4);
            o[2] = o[3] = 0;
        }
    }
}

void refh_trace_tris(const float *tris, int n, int wide, const float *rays, int nrays, int any, float *hits)
{
    void *h = refh_tris_new(tris, n, wide);
    refh_tris_trace(h, rays, nrays, any, hits);
    refh_tris_free(h);
}

/* ---- the path tracer's pieces ---- */
#include "scene/brdf.h"
#include "rt/accel.h"
#include "rt/tracer.h"

/* the BSDF tables the reference was built with (brdf_tables.h): E, B, E_avg, B_avg */
void refh_brdf_tables(float *out)
{
    memcpy(out, BRDF_E, sizeof(BRDF_E));
    memcpy(out + 1024, BRDF_B, sizeof(BRDF_B));
    memcpy(out + 2048, BRDF_E_AVG, sizeof(BRDF_E_AVG));
    memcpy(out + 2080, BRDF_B_AVG, sizeof(BRDF_B_AVG));
}

/* one BSDF case through brdf.h (the layout of lro_bsdf_case): material m = type, albedo.rgb,
 * ior, metallic, roughness, transmission (9 floats) */
void refh_bsdf_case(const float *m, int outside, const float *n, const float *v, const float *l, float ul, float u1,
                    float u2, double *out)
{
    Material mat;
    memset(&mat, 0, sizeof(mat));
    mat.type = (MatType)(int)m[0];
    mat.albedo = vec3(m[1], m[2], m[3]);
    mat.ior = m[4]; mat.metallic = m[5]; mat.roughness = m[6]; mat.transmission = m[7];
    Bsdf b = bsdf_make(&mat, outside);
    Vec3 N = vec3(n[0], n[1], n[2]), V = vec3(v[0], v[1], v[2]), L = vec3(l[0], l[1], l[2]);
    real pdf;
    Vec3 f = bsdf_eval(&b, N, V, L, &pdf);
    out[0] = f.x; out[1] = f.y; out[2] = f.z; out[3] = pdf;
    BsdfSample s;
    memset(&s, 0, sizeof(s));
    bsdf_view(&b, v_dot(V, N));
    int lobe = bsdf_pick(&b, ul);
    int ok = bsdf_sample_lobe(&b, lobe, N, V, u1, u2, &s);
    out[4] = ok; out[5] = lobe;
    out[6] = s.l.x; out[7] = s.l.y; out[8] = s.l.z;
    out[9] = s.w.x; out[10] = s.w.y; out[11] = s.w.z;
    out[12] = s.pdf; out[13] = s.delta; out[14] = s.through; out[15] = 0;
}

/* trace_image (tracer.c:329-356) of a scene of triangles (12 floats each, material in the top
 * 6 bits of word 11) and one quad lamp (q, u, v: 9 floats, material lamp_mat), materials as
 * refh_bsdf_case's 9 floats + emission.rgb (12 floats), built through the reference's own
 * scene API (scene_add_triangle, make_quad, scene_set_accel), camera 12 floats (origin,
 * upper_left, horizontal, vertical); the mean of spp samples per pixel into px (rgb). */
static double refh_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* refh_render, also giving the time trace_image took (the scene's build excluded) */
void refh_render_t(const float *tris, int ntris, const float *lamp, int lamp_mat, const float *mats, int nmats,
                   const float *cam, int W, int H, int spp, uint64_t seed, int max_depth, int mis, double *px,
                   double *trace_ms);

void refh_render(const float *tris, int ntris, const float *lamp, int lamp_mat, const float *mats, int nmats,
                 const float *cam, int W, int H, int spp, uint64_t seed, int max_depth, int mis, double *px)
{
    refh_render_t(tris, ntris, lamp, lamp_mat, mats, nmats, cam, W, H, spp, seed, max_depth, mis, px, NULL);
}

void refh_render_t(const float *tris, int ntris, const float *lamp, int lamp_mat, const float *mats, int nmats,
                   const float *cam, int W, int H, int spp, uint64_t seed, int max_depth, int mis, double *px,
                   double *trace_ms)
{
    Scene s;
    memset(&s, 0, sizeof(s));
    for (int i = 0; i < nmats; i++) {
        const float *m = mats + 12 * (size_t)i;
        Material mat;
        memset(&mat, 0, sizeof(mat));
        mat.type = (MatType)(int)m[0];
        mat.albedo = vec3(m[1], m[2], m[3]);
        mat.ior = m[4]; mat.metallic = m[5]; mat.roughness = m[6]; mat.transmission = m[7];
        mat.emission = vec3(m[8], m[9], m[10]);
        scene_add_material(&s, mat);
    }
    s.light = -1;
    if (lamp)
        s.light = scene_add_prim(&s, make_quad(vec3(lamp[0], lamp[1], lamp[2]), vec3(lamp[3], lamp[4], lamp[5]),
                                               vec3(lamp[6], lamp[7], lamp[8]), lamp_mat));
    for (int i = 0; i < ntris; i++) {
        const float *t = tris + 12 * (size_t)i;
        uint32_t w;
        memcpy(&w, t + 11, 4);
        Vec3 a = vec3(t[0], t[1], t[2]);
        Vec3 b = v_add(a, vec3(t[4], t[5], t[6])), c = v_add(a, vec3(t[8], t[9], t[10]));
        scene_add_triangle(&s, a, b, c, NULL, (int)(w >> 26));
    }
    AccelParams ap = accel_default_params();
    scene_set_accel(&s, ACCEL_BVH, &ap);
    Camera c;
    c.origin = vec3(cam[0], cam[1], cam[2]);
    c.upper_left = vec3(cam[3], cam[4], cam[5]);
    c.horizontal = vec3(cam[6], cam[7], cam[8]);
    c.vertical = vec3(cam[9], cam[10], cam[11]);
    TraceParams tp = trace_default_params();
    tp.max_depth = max_depth;
    tp.mis = mis;
    Vec3 *out = calloc((size_t)W * H, sizeof(Vec3));
    const double t0 = refh_now_ms();
    trace_image(&s, &c, &tp, W, H, spp, seed, 0, out, NULL, NULL);
    if (trace_ms) *trace_ms = refh_now_ms() - t0;
    for (int i = 0; i < W * H; i++) { px[3 * i] = out[i].x; px[3 * i + 1] = out[i].y; px[3 * i + 2] = out[i].z; }
    free(out);
    scene_free(&s);
}

/* frame_render_rt (frame.c:5-12) with light tracing for caustics: the scene of refh_render plus
 * spheres (centre xyz, radius, material: 5 floats each); trace_image's mean of spp samples per
 * pixel with tp.skip_caustics = skip (trace: 0 leaves it out), then, photons > 0,
 * light_trace_caustics with W H spp photons light paths added (its own seed drawn from seed).
 * px: (H, W, 3) doubles. */
void refh_render_lt(const float *tris, int ntris, const float *lamp, int lamp_mat, const float *spheres, int nspheres,
                    const float *mats, int nmats, const float *cam, int W, int H, int spp, uint64_t seed, int max_depth,
                    int trace, int skip, double photons, double *px)
{
    Scene s;
    memset(&s, 0, sizeof(s));
    for (int i = 0; i < nmats; i++) {
        const float *m = mats + 12 * (size_t)i;
        Material mat;
        memset(&mat, 0, sizeof(mat));
        mat.type = (MatType)(int)m[0];
        mat.albedo = vec3(m[1], m[2], m[3]);
        mat.ior = m[4]; mat.metallic = m[5]; mat.roughness = m[6]; mat.transmission = m[7];
        mat.emission = vec3(m[8], m[9], m[10]);
        scene_add_material(&s, mat);
    }
    s.light = -1;
    if (lamp)
        s.light = scene_add_prim(&s, make_quad(vec3(lamp[0], lamp[1], lamp[2]), vec3(lamp[3], lamp[4], lamp[5]),
                                               vec3(lamp[6], lamp[7], lamp[8]), lamp_mat));
    for (int i = 0; i < nspheres; i++) {
        const float *q = spheres + 5 * (size_t)i;
        scene_add_prim(&s, make_sphere(vec3(q[0], q[1], q[2]), q[3], (int)q[4]));
    }
    for (int i = 0; i < ntris; i++) {
        const float *t = tris + 12 * (size_t)i;
        uint32_t w;
        memcpy(&w, t + 11, 4);
        Vec3 a = vec3(t[0], t[1], t[2]);
        Vec3 b = v_add(a, vec3(t[4], t[5], t[6])), c = v_add(a, vec3(t[8], t[9], t[10]));
        scene_add_triangle(&s, a, b, c, NULL, (int)(w >> 26));
    }
    AccelParams ap = accel_default_params();
    scene_set_accel(&s, ACCEL_BVH, &ap);
    Camera c;
    c.origin = vec3(cam[0], cam[1], cam[2]);
    c.upper_left = vec3(cam[3], cam[4], cam[5]);
    c.horizontal = vec3(cam[6], cam[7], cam[8]);
    c.vertical = vec3(cam[9], cam[10], cam[11]);
    TraceParams tp = trace_default_params();
    tp.max_depth = max_depth;
    tp.skip_caustics = skip;
    Vec3 *out = calloc((size_t)W * H, sizeof(Vec3));
    if (trace) trace_image(&s, &c, &tp, W, H, spp, seed, 0, out, NULL, NULL);
    if (photons > 0)
        light_trace_caustics(&s, &c, W, H, (long long)((double)W * H * spp * photons), &tp, seed ^ 0xC0FFEE1234ULL, out);
    for (int i = 0; i < W * H; i++) { px[3 * i] = out[i].x; px[3 * i + 1] = out[i].y; px[3 * i + 2] = out[i].z; }
    free(out);
    scene_free(&s);
}

/* ---- the reference's sky (engine/scene/env.c), for pinning the oracle's ---- */
#include "scene/env.h"

void *refh_env_new(const float *rgb, int w, int h, float scale, float rot)
{
    float *c = malloc(sizeof(float) * 3 * (size_t)w * h);
    if (!c) return NULL;
    memcpy(c, rgb, sizeof(float) * 3 * (size_t)w * h);
    Env *e = env_from_rgb(c, w, h);
    if (!e) return NULL;
    e->scale = scale;
    env_set_rot(e, rot);
    return e;
}
void refh_env_free(void *e) { env_free((Env *)e); }
void refh_env_tables(void *ev, float *marg, float *cond, float *prob, double *sum)
{
    Env *e = (Env *)ev;
    memcpy(marg, e->marg, sizeof(float) * (size_t)(e->h + 1));
    memcpy(cond, e->cond, sizeof(float) * (size_t)e->h * (e->w + 1));
    for (int y = 0; y < e->h; y++)
        for (int x = 0; x < e->w; x++) prob[(size_t)y * e->w + x] = (float)env_pixel_prob(e, x, y);
    *sum = e->sum;
}
void refh_env_eval(void *e, const float *d, float *out)
{
    Vec3 c = env_eval((Env *)e, vec3(d[0], d[1], d[2]));
    out[0] = (float)c.x; out[1] = (float)c.y; out[2] = (float)c.z;
}
void refh_env_sample(void *e, float u1, float u2, float *d, float *pdf, float *le)
{
    Vec3 dd = vec3(0, 0, 0);
    real p = 0;
    Vec3 c = env_sample((Env *)e, u1, u2, &dd, &p);
    d[0] = (float)dd.x; d[1] = (float)dd.y; d[2] = (float)dd.z;
    *pdf = (float)p;
    le[0] = (float)c.x; le[1] = (float)c.y; le[2] = (float)c.z;
}
float refh_env_pdf(void *e, const float *d) { return (float)env_pdf((Env *)e, vec3(d[0], d[1], d[2])); }
void refh_sky(int w, float elev, float azim, float *rgb)
{
    Env *e = env_procedural_sky(w, elev, azim);
    if (!e) return;
    memcpy(rgb, e->rgb, sizeof(float) * 3 * (size_t)e->w * e->h);
    env_free(e);
}
