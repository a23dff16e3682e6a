/* dropin_harness.c - TEST INFRASTRUCTURE ONLY (built into oracle/_ref/libref_dropin.so with the
 * reference's own CPU engine and little_renderer_amd/binding/gpu_frame_hip.c, linked against
 * liblr_hip.so). It drives the binding the way the reference's callers do (apps/render.c:100-147,
 * tests/test_gpu_trace.c:59-71): the reference's own scenes (scene_cornell_box, scene_add_instance,
 * env_procedural_sky), its camera_make and Image, gpu_frame_open / gpu_frame_trace /
 * gpu_frame_close; and beside each GPU result the reference's CPU one on the same scene
 * (trace_image with bin/render's seed 0, scene_intersect, scene_occluded). Plain arrays out, for
 * the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lr.h"
#include "app/gpu_frame.h"
#include "rt/accel.h"
#include "rt/tracer.h"
#include "scene/env.h"
#include "base/rng.h"

#include "../little_renderer_amd/binding/gpu_frame_hip.h"

/* scene 0: the Cornell box with its two boxes (SCENE_BOXES); 1: with a mirror and a glass ball
 * (SCENE_SPHERES); 2: instances under the procedural sky: unit balls and boxes of several
 * materials on a floor box, and a mesh (a box's 12 triangles) twice */
static int make_scene(Scene *s, int kind)
{
    memset(s, 0, sizeof(*s));
    if (kind <= 1) {
        SceneOptions so;
        memset(&so, 0, sizeof(so));
        so.kind = kind == 0 ? SCENE_BOXES : SCENE_SPHERES;
        if (scene_cornell_box(s, &so)) return 1;
        AccelParams ap = accel_default_params();
        scene_set_accel(s, ACCEL_BVH, &ap);
        return 0;
    }
    int white = scene_add_material(s, material_diffuse(vec3(0.7, 0.7, 0.7)));
    int gold = scene_add_material(s, material_pbr(vec3(1.0, 0.78, 0.34), 1, 0.3, 1.5, 0));
    int blue = scene_add_material(s, material_pbr(vec3(0.1, 0.25, 0.7), 0, 0.4, 1.5, 0));
    int glass = scene_add_material(s, material_glass(vec3(1, 1, 1), 1.5));
    int red = scene_add_material(s, material_diffuse(vec3(0.7, 0.1, 0.1)));
    s->light = -1;
    /* a mesh: the box [-1, 1]^3 as 12 triangles (flat) */
    static const int f[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
    Vec3 c[8];
    for (int i = 0; i < 8; i++) c[i] = vec3(i & 4 ? 1 : -1, i & 2 ? 1 : -1, i & 1 ? 1 : -1);
    Triangle tr[12];
    memset(tr, 0, sizeof(tr));
    for (int k = 0; k < 6; k++)
        for (int h = 0; h < 2; h++) {
            Triangle *t = &tr[2 * k + h];
            Vec3 a = c[f[k][0]], b = c[f[k][1 + h]], d = c[f[k][2 + h]];
            t->q = a;
            t->u = v_sub(b, a);
            t->v = v_sub(d, a);
            t->n[0] = t->n[1] = t->n[2] = -1;
            t->t[0] = t->t[1] = t->t[2] = -1;
            t->material = red;
        }
    Triangle *copy = malloc(sizeof(tr));
    memcpy(copy, tr, sizeof(tr));
    int mesh = scene_add_mesh_arrays(s, copy, 12, NULL, 0, NULL, NULL, 0);
    if (mesh < 0) return 1;
    Mat3 id = mat3_identity();
    scene_add_instance(s, INST_BOX, xform_axes(id, vec3(4, 0.1, 4), vec3(0, -0.1, 0)), white);
    scene_add_instance(s, INST_SPHERE, xform_axes(id, vec3(0.5, 0.5, 0.5), vec3(-1.2, 0.5, 0)), gold);
    scene_add_instance(s, INST_SPHERE, xform_axes(id, vec3(0.5, 0.5, 0.5), vec3(0, 0.5, 0.3)), glass);
    scene_add_instance(s, INST_BOX, xform_axes(id, vec3(0.4, 0.4, 0.4), vec3(1.2, 0.4, -0.2)), blue);
    scene_add_instance(s, mesh, xform_rot_y(vec3(0.6, 0.3, -1.2), 30, 0.3), -1);
    scene_add_instance(s, mesh, xform_rot_y(vec3(-0.5, 0.25, -1.0), -20, 0.25), white);
    Env *e = env_procedural_sky(256, 35, -30);
    if (!e) return 1;
    e->scale = 1;
    env_set_rot(e, 0.3f);
    s->env = e;
    s->cam_from = vec3(0.3, 1.6, -4.2);
    s->cam_at = vec3(0, 0.4, 0);
    s->cam_up = vec3(0, 1, 0);
    s->cam_vfov_deg = 40;
    AccelParams ap = accel_default_params();
    scene_set_accel(s, ACCEL_BVH, &ap);
    return 0;
}

static void to_floats(const Image *img, float *out)
{
    for (int i = 0; i < img->w * img->h; i++) {
        out[3 * i] = (float)img->px[i].x;
        out[3 * i + 1] = (float)img->px[i].y;
        out[3 * i + 2] = (float)img->px[i].z;
    }
}

/* trace_image; photons > 0: with skip_caustics, and light_trace_caustics' W H spp photons light
 * paths added (frame_render_rt's way, engine/app/frame.c:5-12; the light paths' seed the
 * harness's own) */
static void cpu_image(const Scene *s, const Camera *cam, int W, int H, int spp, uint64_t seed, double photons, float *out)
{
    TraceParams tp = trace_default_params();
    tp.max_depth = 8;
    tp.mis = 1;
    tp.skip_caustics = photons > 0;
    Vec3 *px = malloc(sizeof(Vec3) * (size_t)W * H);
    trace_image(s, cam, &tp, W, H, spp, seed, 0, px, NULL, NULL);
    if (photons > 0)
        light_trace_caustics(s, cam, W, H, (long long)((double)W * H * spp * photons), &tp, seed * 7919 + 17, px);
    for (int i = 0; i < W * H; i++) {
        out[3 * i] = (float)px[i].x;
        out[3 * i + 1] = (float)px[i].y;
        out[3 * i + 2] = (float)px[i].z;
    }
    free(px);
}

/* The scene's image, W x H at spp samples (photons > 0: with light-traced caustics, that many
 * light paths a pixel and sample), by the binding (gpu: 3 floats a pixel, traced twice
 * on a new GPU and view each time, as two runs of bin/render: gpu2 the second) and by the
 * reference's trace_image, K times with seeds 1 .. K (cpu: K images, independent of the GPU's
 * samples, whose spread is each pixel's Monte Carlo noise). move = 1 (the instanced scene): after
 * the first images, two instances move (scene_move_instance, scene_update_instances) and the same
 * view traces again into gpu2, the CPU the moved scene. 0, 1 if the binding refused or failed,
 * -1 no scene. */
int dropin_render(int kind, int W, int H, int spp, int move, int K, double photons, float *gpu, float *gpu2, float *cpu,
                  int *passes)
{
    static Scene s;
    if (make_scene(&s, kind)) return -1;
    Camera cam = camera_make(s.cam_from, s.cam_at, s.cam_up, s.cam_vfov_deg, (real)W / H);
    Image img = image_create(W, H);
    int rc = 0;
    for (int run = 0; run < 2 && !rc; run++) {
        Gpu *g;
        GpuView *v;
        if (gpu_frame_open(W, H, &g, &v)) {
            rc = 1;
            break;
        }
        rc = gpu_frame_trace(g, v, &s, &cam, NULL, NULL, spp, 8, 1, 0, 1, photons, &img, 0);
        if (!rc) {
            to_floats(&img, run == 0 ? gpu : gpu2);
            passes[run] = gpu_frame_passes(v);
        }
        if (!rc && move && run == 1) {
            scene_move_instance(&s, 1, xform_axes(mat3_identity(), vec3(0.5, 0.5, 0.5), vec3(-0.9, 0.5, 0.8)));
            scene_move_instance(&s, 3, xform_axes(mat3_identity(), vec3(0.3, 0.6, 0.3), vec3(1.0, 0.6, 0.5)));
            scene_update_instances(&s, NULL);
            rc = gpu_frame_trace(g, v, &s, &cam, NULL, NULL, spp, 8, 1, 0, 1, photons, &img, 0);
            if (!rc) to_floats(&img, gpu2);
        }
        gpu_frame_close(g, v);
    }
    for (int k = 0; k < K && !rc; k++)
        cpu_image(&s, &cam, W, H, spp, 1 + (uint64_t)k, photons, cpu + 3 * (size_t)W * H * k);
    image_free(&img);
    scene_free(&s);
    return rc;
}

/* n random rays from inside the scene's bounds (PCG32 seed): the binding's batch queries and the
 * CPU's, 10 floats a ray: gpu found, prim, inst, t, front_face, cpu found, prim, inst, t,
 * front_face; then occluded (gpu, cpu) over (tmin, tmax / 2), 2 floats a ray in occ. */
int dropin_queries(int kind, int n, uint64_t seed, float *out, float *occ)
{
    static Scene s;
    if (make_scene(&s, kind)) return -1;
    Rng rng;
    rng_seed(&rng, seed, 9);
    Vec3 lo = kind <= 1 ? vec3(10, 10, 10) : vec3(-2, 0.05, -2), hi = kind <= 1 ? vec3(545, 545, 545) : vec3(2, 2, 2);
    Ray *rays = malloc(sizeof(Ray) * (size_t)n);
    for (int i = 0; i < n; i++) {
        Vec3 o = vec3(lo.x + (hi.x - lo.x) * rng_real(&rng), lo.y + (hi.y - lo.y) * rng_real(&rng),
                      lo.z + (hi.z - lo.z) * rng_real(&rng));
        real z = 2 * rng_real(&rng) - 1, ph = 2 * PI * rng_real(&rng), r = r_sqrt(r_fmax(0, 1 - z * z));
        rays[i].o = o;
        rays[i].d = vec3(r * r_cos(ph), z, r * r_sin(ph));
    }
    const real tmin = 2e-3f, tmax = 1e30f;
    Hit *hg = calloc((size_t)n, sizeof(Hit));
    int *fg = calloc((size_t)n, sizeof(int)), *og = calloc((size_t)n, sizeof(int));
    Gpu *g;
    GpuView *v;
    int rc = gpu_frame_open(64, 64, &g, &v);
    if (!rc) {
        rc = gpu_scene_intersect(v, &s, rays, n, tmin, tmax, hg, fg);
        if (!rc) rc = gpu_scene_occluded(v, &s, rays, n, tmin, kind <= 1 ? 200 : 1.0f, og);
        gpu_frame_close(g, v);
    }
    for (int i = 0; i < n && !rc; i++) {
        Hit hc;
        int fc = scene_intersect(&s, &rays[i], tmin, tmax, &hc);
        float *o = out + 10 * (size_t)i;
        o[0] = (float)fg[i];
        o[1] = fg[i] ? (float)hg[i].prim : -1;
        o[2] = fg[i] ? (float)hg[i].inst : -1;
        o[3] = fg[i] ? (float)hg[i].t : 0;
        o[4] = fg[i] ? (float)hg[i].front_face : 0;
        o[5] = (float)fc;
        o[6] = fc ? (float)hc.prim : -1;
        o[7] = fc ? (float)hc.inst : -1;
        o[8] = fc ? (float)hc.t : 0;
        o[9] = fc ? (float)hc.front_face : 0;
        occ[2 * i] = (float)og[i];
        occ[2 * i + 1] = (float)scene_occluded(&s, &rays[i], tmin, kind <= 1 ? 200 : 1.0f);
    }
    free(rays);
    free(hg);
    free(fg);
    free(og);
    scene_free(&s);
    return rc;
}
