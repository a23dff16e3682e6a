/* gpu_frame_hip.c - little_renderer's offline GPU image (engine/app/gpu_frame.h) and batches of
 * scene ray queries (engine/rt/accel.h:43-45), on liblr_hip.so (include/lr_hip.h).
 *
 * This is the reference side of the boundary: the file a maintainer builds in place of
 * engine/app/gpu_frame.c, in an engine built without its WebGPU layer (NOGPU=1: gpu.c and
 * view_gpu.c are then their no-GPU stubs, which never look inside a Gpu or a GpuView). Callers
 * such as apps/render.c:100-147 keep their code: gpu_frame_open / gpu_frame_trace /
 * gpu_frame_close have the reference's signatures and return conventions (0 on success; 1 when
 * the GPU is not there, cannot hold the scene, or failed: the caller renders on the CPU).
 * It is compiled against the reference's headers (-I engine) and links -llr_hip; it keeps
 * nothing of the reference's GPU code.
 *
 * What the device copy holds: the scene's primitives (spheres, quads), its triangles (flat:
 * Triangle.n ignored only when every index is -1, else the scene is refused), its materials,
 * the lamp (Scene.light) or the sky (Scene.env), or a scene of instances (meshes, unit spheres
 * and boxes) under the sky. It is kept between calls and made again only when the scene
 * changed: by tris_version and the primitives' and materials' contents; an instanced scene's
 * moved instances are sent alone, found by their versions as upload_instances finds them
 * (engine/gpu/trace_gpu.c:1391-1452, scene_changed_run, engine/scene/scene.h:359-363), and a
 * dynamic mesh whose version moved on is sent and its tree rebuilt in place (:1486-1512).
 * Refused (1, the CPU renders): textures, media (volumes), vertex normals, a scene of instances
 * that also has world primitives or triangles. Light-traced caustics (photons > 0) are traced on
 * the device: photons x w x h light paths a pass (the GPU's caustics pass, trace_gpu.c:1587-1594),
 * the path tracer leaving their paths out (skip_caustics).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "app/gpu_frame.h"
#include "rt/accel.h"
#include "rt/intersect.h"
#include "scene/env.h"

#include "lr_hip.h"
#include "lr_scene.h"
#include "gpu_frame_hip.h"

/* the ray offsets of the float build's path tracer (RT_EPS, RT_SHADOW_MARGIN,
 * engine/rt/tracer.c:15-23): the device traces in float whatever `real` is */
#define HIP_RT_EPS 2e-3f
#define HIP_RT_MARGIN 2e-2f

typedef struct {
    LrhContext *ctx;
} HipGpu;

typedef struct {
    HipGpu *g;
    int w, h;
    int passes;                 /* samples per pixel of the last image */
    /* the device copy and what it was made from */
    LrhPtScene *ps;
    const Scene *of;
    int instanced;
    unsigned tris_version, inst_version, insts_id;
    int nprims, ntris, ninsts, nmeshes, light;
    const Env *env;
    unsigned env_id;
    float *prims;               /* nprims x 24: the records it was made with */
    LrhMaterial mats[MAX_MATS];
    int nmats;
    unsigned *mesh_version;     /* nmeshes */
    int *mesh_first;            /* nmeshes + 1: each mesh's first triangle in the device copy */
    int depth;
} HipView;

static void say(const char *what, LrhContext *c)
{
    (void)c;
    fprintf(stderr, "gpu (HIP): %s: %s\n", what, lrh_last_error());
}

/* ---- the scene's device copy ---------------------------------------------------------------- */

static void put_material(LrhMaterial *o, const Material *m)
{
    memset(o, 0, sizeof(*o));
    o->type = (int)m->type;
    o->albedo[0] = (float)m->albedo.x;
    o->albedo[1] = (float)m->albedo.y;
    o->albedo[2] = (float)m->albedo.z;
    o->emission[0] = (float)m->emission.x;
    o->emission[1] = (float)m->emission.y;
    o->emission[2] = (float)m->emission.z;
    o->ior = (float)m->ior;
    o->metallic = (float)m->metallic;
    o->roughness = (float)m->roughness;
    o->transmission = (float)m->transmission;
}

/* a primitive as lr_hip.h's 24-float record, from the fields make_sphere / make_quad computed
 * (engine/scene/scene.c:16-40), so the device tests the shape the CPU tests */
static void put_prim(float *o, const Primitive *p)
{
    memset(o, 0, 24 * sizeof(float));
    if (p->type == PRIM_SPHERE) {
        o[0] = (float)p->center.x;
        o[1] = (float)p->center.y;
        o[2] = (float)p->center.z;
        o[3] = (float)p->radius;
        o[7] = 0.0f;
        o[11] = (float)p->material;
        return;
    }
    const Vec3 *v[5] = {&p->q, &p->u, &p->v, &p->normal, &p->w};
    for (int k = 0; k < 5; k++) {
        o[4 + 4 * k] = (float)v[k]->x;
        o[5 + 4 * k] = (float)v[k]->y;
        o[6 + 4 * k] = (float)v[k]->z;
    }
    o[7] = 1.0f;
    o[11] = (float)p->material;
    o[15] = (float)p->area;
    o[19] = (float)p->plane_d;
}

/* a triangle in lr_hip.h's 48-byte layout (flat: no normal indices), its material in 6 bits */
static void put_tri(float *o, const Triangle *t)
{
    const Vec3 *e[3] = {&t->q, &t->u, &t->v};
    for (int k = 0; k < 3; k++) {
        o[4 * k] = (float)e[k]->x;
        o[4 * k + 1] = (float)e[k]->y;
        o[4 * k + 2] = (float)e[k]->z;
    }
    const uint32_t w[3] = {0xffffffffu, 0u, (uint32_t)(t->material & 63) << 26};
    for (int k = 0; k < 3; k++) memcpy(&o[4 * k + 3], &w[k], 4);
}

static void put_instance(LrhInstance *o, const Instance *in)
{
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) o->xform[4 * r + c] = (float)in->xf.m[r][c];
    o->kind = in->mesh >= 0 ? in->mesh : (in->mesh == INST_SPHERE ? LRH_INST_SPHERE : LRH_INST_BOX);
    o->material = in->material;
}

/* why the device cannot hold the scene, or NULL */
static const char *refusal(const Scene *s)
{
    if (s->nvolumes > 0) return "media (volumes) are traced on the CPU";
    if (s->nmats > MAX_MATS) return "too many materials";
    for (int i = 0; i < s->nmats; i++)
        if (material_textured(&s->mats[i])) return "textured materials are traced on the CPU";
    for (int i = 0; i < s->ntris; i++)
        if (s->tris[i].n[0] >= 0 || s->tris[i].n[1] >= 0 || s->tris[i].n[2] >= 0)
            return "vertex normals are traced on the CPU";
    if (s->ninsts > 0 && (s->nprims > 0 || s->ntris > 0))
        return "instances beside world primitives or triangles are traced on the CPU";
    for (int m = 0; m < s->nmeshes; m++)
        for (int i = 0; i < s->meshes[m].ntris; i++) {
            const Triangle *t = &s->meshes[m].tris[i];
            if (t->n[0] >= 0 || t->n[1] >= 0 || t->n[2] >= 0) return "vertex normals are traced on the CPU";
        }
    if (s->ninsts == 0 && s->nprims + s->ntris == 0) return "an empty scene";
    return NULL;
}

static void drop_copy(HipView *v)
{
    if (v->ps) lrh_pt_scene_free(v->ps);
    free(v->prims);
    free(v->mesh_version);
    free(v->mesh_first);
    v->ps = NULL;
    v->prims = NULL;
    v->mesh_version = NULL;
    v->mesh_first = NULL;
    v->of = NULL;
}

/* the copy's sky made the scene's (none: taken away) */
static int sky_of(HipView *v, const Scene *s)
{
    const Env *e = s->env;
    const int had = v->env != NULL;
    v->env = e;
    v->env_id = e ? e->id : 0;
    if (!e) return had ? lrh_pt_scene_sky(v->ps, NULL, 0, 0, 1.0f, 0.0f) : 0;
    return lrh_pt_scene_sky(v->ps, e->rgb, e->w, e->h, (float)e->scale, (float)e->rot);
}

/* a new device copy of a one-level scene (primitives and triangles) */
static int copy_flat(HipView *v, const Scene *s, const float *prims, const LrhMaterial *mats)
{
    float *tr = malloc(((size_t)s->ntris * 12 + 1) * sizeof(float));
    if (!tr) return LRH_ENOMEM;
    for (int i = 0; i < s->ntris; i++) put_tri(tr + 12 * (size_t)i, &s->tris[i]);
    int rc = lrh_pt_scene(v->g->ctx, prims, s->nprims, tr, s->ntris, mats, s->nmats, s->light, HIP_RT_EPS,
                          HIP_RT_MARGIN, LRH_PLOC, &v->ps);
    free(tr);
    return rc;
}

/* a new device copy of a scene of instances */
static int copy_instanced(HipView *v, const Scene *s, const LrhMaterial *mats)
{
    size_t total = 0;
    v->mesh_first = malloc(((size_t)s->nmeshes + 1) * sizeof(int));
    v->mesh_version = malloc(((size_t)s->nmeshes + 1) * sizeof(unsigned));
    if (!v->mesh_first || !v->mesh_version) return LRH_ENOMEM;
    for (int m = 0; m < s->nmeshes; m++) {
        v->mesh_first[m] = (int)total;
        v->mesh_version[m] = s->meshes[m].version;
        total += (size_t)s->meshes[m].ntris;
    }
    v->mesh_first[s->nmeshes] = (int)total;
    float *tr = malloc((total * 12 + 1) * sizeof(float));
    LrhInstance *inst = malloc(((size_t)s->ninsts + 1) * sizeof(LrhInstance));
    if (!tr || !inst) {
        free(tr);
        free(inst);
        return LRH_ENOMEM;
    }
    for (int m = 0; m < s->nmeshes; m++)
        for (int i = 0; i < s->meshes[m].ntris; i++)
            put_tri(tr + 12 * ((size_t)v->mesh_first[m] + i), &s->meshes[m].tris[i]);
    for (int i = 0; i < s->ninsts; i++) put_instance(&inst[i], &s->insts[i]);
    int rc = lrh_pt_scene_instanced(v->g->ctx, tr, v->mesh_first, s->nmeshes, mats, s->nmats, inst, s->ninsts,
                                    HIP_RT_EPS, LRH_PLOC, &v->ps);
    free(tr);
    free(inst);
    return rc;
}

/* An instanced copy brought up to date: the instances changed since its version, in the runs
 * scene_changed_run gives (gaps of up to 8 unchanged ones joined, as upload_instances sends
 * them), then the dynamic meshes whose version moved on. 0, or < 0 (then made anew). */
static int update_instanced(HipView *v, const Scene *s)
{
    if (s->insts_id != v->insts_id || s->ninsts != v->ninsts || s->nmeshes != v->nmeshes) return -1;
    for (int m = 0; m < s->nmeshes; m++)
        if (s->meshes[m].ntris != v->mesh_first[m + 1] - v->mesh_first[m]) return -1;
    int rc = 0;
    if (s->inst_version != v->inst_version) {
        int *idx = malloc(((size_t)s->ninsts + 1) * sizeof(int));
        LrhInstance *inst = malloc(((size_t)s->ninsts + 1) * sizeof(LrhInstance));
        if (!idx || !inst) {
            free(idx);
            free(inst);
            return -1;
        }
        int n = 0;
        for (int i = 0, a, b; scene_changed_run(s, v->inst_version, 8, &i, &a, &b);)
            for (int k = a; k <= b; k++) {
                idx[n] = k;
                put_instance(&inst[n], &s->insts[k]);
                n++;
            }
        rc = lrh_pt_scene_move(v->ps, idx, inst, n, LRH_PLOC, NULL);
        free(idx);
        free(inst);
        if (rc) return rc;
        v->inst_version = s->inst_version;
    }
    for (int m = 0; m < s->nmeshes; m++) {
        const SceneMesh *sm = &s->meshes[m];
        if (sm->version == v->mesh_version[m]) continue;
        float *tr = malloc(((size_t)sm->ntris * 12 + 1) * sizeof(float));
        if (!tr) return -1;
        for (int i = 0; i < sm->ntris; i++) put_tri(tr + 12 * (size_t)i, &sm->tris[i]);
        rc = lrh_pt_scene_mesh(v->ps, m, tr, sm->ntris, LRH_PLOC, NULL);
        free(tr);
        if (rc) return rc;
        v->mesh_version[m] = sm->version;
    }
    return 0;
}

/* the device copy of s, made or brought up to date; NULL if the device cannot hold it */
static LrhPtScene *scene_copy(HipView *v, const Scene *s)
{
    const char *why = refusal(s);
    if (why) {
        fprintf(stderr, "gpu (HIP): %s\n", why);
        return NULL;
    }
    LrhMaterial mats[MAX_MATS];
    for (int i = 0; i < s->nmats; i++) put_material(&mats[i], &s->mats[i]);
    float *prims = malloc(((size_t)s->nprims * 24 + 1) * sizeof(float));
    if (!prims) return NULL;
    for (int i = 0; i < s->nprims; i++) put_prim(prims + 24 * (size_t)i, &s->prims[i]);
    const int instanced = s->ninsts > 0;
    const int same_base = v->ps && v->of == s && v->instanced == instanced && v->nmats == s->nmats &&
                          !memcmp(v->mats, mats, sizeof(LrhMaterial) * (size_t)s->nmats) && v->light == s->light;
    int keep = 0;
    if (same_base && !instanced)
        keep = v->tris_version == s->tris_version && v->ntris == s->ntris && v->nprims == s->nprims &&
               !memcmp(v->prims, prims, sizeof(float) * 24 * (size_t)s->nprims);
    else if (same_base && instanced)
        keep = update_instanced(v, s) == 0;
    if (keep) {
        free(prims);
        if ((s->env != v->env || (s->env && s->env->id != v->env_id)) && sky_of(v, s)) {
            say("the sky", v->g->ctx);
            drop_copy(v);
            return NULL;
        }
        return v->ps;
    }
    drop_copy(v);
    v->env = NULL;
    v->instanced = instanced;
    int rc = instanced ? copy_instanced(v, s, mats) : copy_flat(v, s, prims, mats);
    if (!rc) rc = sky_of(v, s);
    if (rc) {
        say(rc == LRH_EREFUSED ? "the scene is not one it can hold" : "the scene's copy", v->g->ctx);
        free(prims);
        drop_copy(v);
        return NULL;
    }
    v->prims = prims;
    v->of = s;
    memcpy(v->mats, mats, sizeof(LrhMaterial) * (size_t)s->nmats);
    v->nmats = s->nmats;
    v->light = s->light;
    v->tris_version = s->tris_version;
    v->ntris = s->ntris;
    v->nprims = s->nprims;
    v->ninsts = s->ninsts;
    v->nmeshes = s->nmeshes;
    v->inst_version = s->inst_version;
    v->insts_id = s->insts_id;
    int nshapes = 0;
    size_t bytes = 0;
    lrh_pt_scene_info(v->ps, &v->depth, &nshapes, &bytes);
    return v->ps;
}

/* ---- gpu_frame.h ------------------------------------------------------------------------------ */

int gpu_frame_open(int w, int h, Gpu **gpu, GpuView **view)
{
    *gpu = NULL;
    *view = NULL;
    if (w <= 0 || h <= 0) return 1;
    HipGpu *g = calloc(1, sizeof(HipGpu));
    HipView *v = calloc(1, sizeof(HipView));
    if (!g || !v || lrh_create(&g->ctx, 0, NULL) != 0) {
        if (g && !g->ctx) fprintf(stderr, "gpu (HIP): %s\n", lrh_last_error());
        free(g);
        free(v);
        return 1;
    }
    v->g = g;
    v->w = w;
    v->h = h;
    *gpu = (Gpu *)g;
    *view = (GpuView *)v;
    return 0;
}

void gpu_frame_close(Gpu *gpu, GpuView *view)
{
    HipView *v = (HipView *)view;
    HipGpu *g = (HipGpu *)gpu;
    if (v) {
        drop_copy(v);
        free(v);
    }
    if (g) {
        lrh_destroy(g->ctx);
        free(g);
    }
}

/* camera_make's four vectors (engine/scene/camera.h:14-19) as lr_hip.h's 12 floats */
static void put_camera(float c[12], const Camera *cam)
{
    const Vec3 *v[4] = {&cam->origin, &cam->upper_left, &cam->horizontal, &cam->vertical};
    for (int k = 0; k < 4; k++) {
        c[3 * k] = (float)v[k]->x;
        c[3 * k + 1] = (float)v[k]->y;
        c[3 * k + 2] = (float)v[k]->z;
    }
}

int gpu_frame_trace(Gpu *gpu, GpuView *view, const Scene *s, const Camera *cam, const RenderParams *rp,
                    const MarchParams *mp, int spp, int max_depth, int mis, real clamp, int sampler, real photons,
                    Image *img, int progress)
{
    (void)rp;
    (void)mp;
    (void)sampler;    /* the device's numbers are the Sobol sampler's (the same expectation) */
    HipGpu *g = (HipGpu *)gpu;
    HipView *v = (HipView *)view;
    if (!g || !v || !s || !cam || !img || img->w <= 0 || img->h <= 0 || spp < 0) return 1;
    LrhPtScene *ps = scene_copy(v, s);
    if (!ps) return 1;
    const int W = img->w, H = img->h, n = W * H;
    float c[12];
    put_camera(c, cam);
    void *d_acc = NULL;
    float *h_acc = malloc((size_t)n * 4 * sizeof(float));
    int rc = h_acc ? lrh_buffer_alloc(g->ctx, (size_t)n * 4 * sizeof(float), &d_acc) : LRH_ENOMEM;
    /* seed 0: trace_image's for bin/render (apps/render.c:80), so -gpu on and off give the
       same samples; the passes in one call (the engine batches them) */
    LrhJob job;
    memset(&job, 0, sizeof(job));
    job.scene = ps;
    memcpy(job.cam, c, sizeof(job.cam));
    job.w = W;
    job.h = H;
    job.spp = spp;
    job.max_depth = max_depth;
    job.mis = mis;
    job.clamp = (float)clamp;
    job.photons = photons > 0 ? (float)photons : 0.0f;
    if (!rc) rc = lrh_pt_render_job(g->ctx, &job, (float *)d_acc);
    if (!rc) rc = lrh_buffer_read(g->ctx, h_acc, d_acc, (size_t)n * 4 * sizeof(float));
    if (d_acc) lrh_buffer_free(g->ctx, d_acc);
    if (rc) {
        say("the image", g->ctx);
        free(h_acc);
        return 1;
    }
    /* wf_resolve's sums and counts to the mean (gpu_frame.c's on_samples) */
    for (int i = 0; i < n; i++) {
        const float k = h_acc[4 * i + 3];
        img->px[i] = k > 0 ? vec3(h_acc[4 * i] / k, h_acc[4 * i + 1] / k, h_acc[4 * i + 2] / k) : vec3(0, 0, 0);
    }
    free(h_acc);
    v->passes = spp;
    if (progress) fprintf(stderr, "\r  100%%\n");
    return 0;
}

int gpu_frame_raster(Gpu *gpu, GpuView *view, const Scene *s, const RenderParams *rp, const MarchParams *mp,
                     int volumes, Framebuffer *fb)
{
    (void)gpu;
    (void)view;
    (void)s;
    (void)rp;
    (void)mp;
    (void)volumes;
    (void)fb;
    return 1;   /* the rasterizer is not this library's: the caller draws on the CPU */
}

int gpu_frame_passes(const GpuView *view)
{
    return view ? ((const HipView *)view)->passes : 0;
}

/* ---- batches of scene ray queries --------------------------------------------------------------- */

/* rays as lr_hip.h's 8 floats (o.xyz tmin | d.xyz tmax) in a device buffer */
static int rays_to_device(HipGpu *g, const Ray *rays, int n, real tmin, real tmax, void **d_rays)
{
    float *r = malloc((size_t)n * 8 * sizeof(float));
    if (!r) return LRH_ENOMEM;
    for (int i = 0; i < n; i++) {
        float *o = r + 8 * (size_t)i;
        o[0] = (float)rays[i].o.x;
        o[1] = (float)rays[i].o.y;
        o[2] = (float)rays[i].o.z;
        o[3] = (float)tmin;
        o[4] = (float)rays[i].d.x;
        o[5] = (float)rays[i].d.y;
        o[6] = (float)rays[i].d.z;
        o[7] = (float)tmax;
    }
    int rc = lrh_buffer_alloc(g->ctx, (size_t)n * 8 * sizeof(float), d_rays);
    if (!rc) rc = lrh_buffer_write(g->ctx, *d_rays, r, (size_t)n * 8 * sizeof(float));
    free(r);
    return rc;
}

static int query(HipView *v, const Scene *s, const Ray *rays, int n, real tmin, real tmax, int any, float **hits)
{
    *hits = NULL;
    LrhPtScene *ps = scene_copy(v, s);
    if (!ps) return 1;
    void *d_rays = NULL, *d_hits = NULL;
    float *h = malloc((size_t)n * 4 * sizeof(float) + 16);
    int rc = h ? rays_to_device(v->g, rays, n, tmin, tmax, &d_rays) : LRH_ENOMEM;
    if (!rc) rc = lrh_buffer_alloc(v->g->ctx, (size_t)n * 4 * sizeof(float), &d_hits);
    if (!rc) rc = lrh_pt_scene_trace(v->g->ctx, ps, (const float *)d_rays, n, any, (float *)d_hits);
    if (!rc) rc = lrh_buffer_read(v->g->ctx, h, d_hits, (size_t)n * 4 * sizeof(float));
    if (d_rays) lrh_buffer_free(v->g->ctx, d_rays);
    if (d_hits) lrh_buffer_free(v->g->ctx, d_hits);
    if (rc) {
        say("the ray queries", v->g->ctx);
        free(h);
        return 1;
    }
    *hits = h;
    return 0;
}

/* the hit record of set_hit (engine/rt/intersect.c:4-10) for a shape the device found: the
 * point, the geometric normal facing the ray, and where on the shape */
static void fill_hit(const HipView *v, const Scene *s, const Ray *ray, float t, uint32_t id, uint32_t inst, float b1,
                     float b2, Hit *h)
{
    memset(h, 0, sizeof(*h));
    h->t = t;
    h->p = v_add(ray->o, v_scale(ray->d, t));
    h->inst = -1;
    Vec3 outward;
    if (v->instanced) {
        const Instance *in = &s->insts[inst];
        const Vec3 od = xform_dir(&in->inv, ray->d), oo = xform_point(&in->inv, ray->o);
        const Vec3 op = v_add(oo, v_scale(od, t));
        Vec3 no;
        h->prim = -1;   /* a shape's (intersect_instances, engine/rt/accel.c:121-130) */
        if (in->mesh == INST_SPHERE) no = op;
        else if (in->mesh == INST_BOX) {
            const real ax = r_fabs(op.x), ay = r_fabs(op.y), az = r_fabs(op.z);
            no = ax >= ay && ax >= az ? vec3(op.x > 0 ? 1 : -1, 0, 0)
                                      : (ay >= az ? vec3(0, op.y > 0 ? 1 : -1, 0) : vec3(0, 0, op.z > 0 ? 1 : -1));
        } else {
            /* (the device's hit carries the instance where a flat scene's carries b1: the
               barycentrics again from the object-space point, as for a quad) */
            const int local = (int)id - v->mesh_first[in->mesh];
            const Triangle *tr = &s->meshes[in->mesh].tris[local];
            no = v_cross(tr->u, tr->v);
            const real nn = v_dot(no, no);
            const Vec3 rel = v_sub(op, tr->q);
            h->prim = local;
            h->uv = nn > 0 ? vec2(v_dot(no, v_cross(rel, tr->v)) / nn, v_dot(no, v_cross(tr->u, rel)) / nn) : vec2(0, 0);
        }
        outward = xform_normal(&in->inv, no);
        h->inst = (int)inst;
    } else if ((int)id < s->nprims) {
        const Primitive *p = &s->prims[id];
        if (p->type == PRIM_SPHERE) outward = v_scale(v_sub(h->p, p->center), 1 / p->radius);
        else {
            const Vec3 rel = v_sub(h->p, p->q), pv = v_cross(rel, p->v), up = v_cross(p->u, rel);
            outward = p->normal;
            h->uv = vec2(v_dot(p->w, pv), v_dot(p->w, up));
        }
        h->prim = (int)id;
    } else {
        const Triangle *tr = &s->tris[id - (uint32_t)s->nprims];
        outward = v_norm(v_cross(tr->u, tr->v));
        h->prim = (int)id;
        h->uv = vec2(b1, b2);
    }
    h->front_face = v_dot(ray->d, outward) < 0;
    h->ng = h->front_face ? outward : v_neg(outward);
    h->n = h->ng;
}

int gpu_scene_intersect(GpuView *view, const Scene *s, const Ray *rays, int n, real tmin, real tmax, Hit *hits,
                        int *found)
{
    HipView *v = (HipView *)view;
    if (!v || !s || n < 0 || (n > 0 && (!rays || !hits || !found))) return 1;
    if (n == 0) return 0;
    float *h = NULL;
    if (query(v, s, rays, n, tmin, tmax, 0, &h)) return 1;
    for (int i = 0; i < n; i++) {
        uint32_t id, inst;
        memcpy(&id, &h[4 * (size_t)i + 1], 4);
        memcpy(&inst, &h[4 * (size_t)i + 2], 4);
        found[i] = id != 0xffffffffu;
        if (found[i])
            fill_hit(v, s, &rays[i], h[4 * (size_t)i], id, inst, h[4 * (size_t)i + 2], h[4 * (size_t)i + 3], &hits[i]);
    }
    free(h);
    return 0;
}

int gpu_scene_occluded(GpuView *view, const Scene *s, const Ray *rays, int n, real tmin, real tmax, int *occluded)
{
    HipView *v = (HipView *)view;
    if (!v || !s || n < 0 || (n > 0 && (!rays || !occluded))) return 1;
    if (n == 0) return 0;
    float *h = NULL;
    if (query(v, s, rays, n, tmin, tmax, 1, &h)) return 1;
    for (int i = 0; i < n; i++) {
        uint32_t id;
        memcpy(&id, &h[4 * (size_t)i + 1], 4);
        occluded[i] = id != 0xffffffffu;
    }
    free(h);
    return 0;
}
