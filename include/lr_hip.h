/* lr_hip.h - the C ABI of little_renderer's path-tracing hot path on AMD MI355X (gfx950).
 *
 * Plain pointers and sizes only: no HIP, torch or WebGPU types. Pointers named d_* are device
 * (HBM) addresses; everything else is host memory. Every call returns 0 on success or a
 * negative LRH_E* code, with a message in lrh_last_error() (per host thread).
 *
 * Layouts are the reference's GPU layouts (engine/gpu/trace_gpu.c:825-911):
 *   box   8 floats  lo.x lo.y lo.z 0 | hi.x hi.y hi.z 0             (build_queue, :587-641)
 *   node  8 words   lo.xyz (f32) a (u32) | hi.xyz (f32) b (u32)     (trace.wgsl:2)
 *                   inner: a, b = absolute child indices; leaf: a = first aux slot,
 *                   b = 0x80000000 | count
 *   tri   12 words  q.xyz n0 | u.xyz n1 | v.xyz (material << 26 | n2) (put_tris, :898-911)
 *   ray   8 floats  o.xyz tmin | d.xyz tmax
 *   hit   4 words   t (f32), shape id (u32, 0xffffffff = miss), b1, b2 (f32)
 */
#ifndef LR_HIP_H
#define LR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LRH_OK 0
#define LRH_EINVAL (-1)     /* a bad argument */
#define LRH_ENOMEM (-2)     /* device memory could not be had */
#define LRH_EDEVICE (-3)    /* the device failed (a lost device: the caller falls back to the CPU) */
#define LRH_EREFUSED (-4)   /* the request does not fit (as gpu_trace_sync's 0, trace_gpu.c:1571-1581) */

#define LRH_LBVH 0
#define LRH_PLOC 1
/* an instanced scene's top level only: the same tree over the instances' new boxes while its cost
 * stays within 1.1 of its last build's, else built again with that build's builder (LR_TLAS_REFIT,
 * engine/rt/accel.c:152-173) */
#define LRH_REFIT 2

typedef struct LrhContext LrhContext;

typedef struct {
    int depth;       /* the tree's depth in levels (check_tree's, trace_gpu.c:768-805) */
    int rounds;      /* PLOC rounds (0 for an LBVH) */
    float ms;        /* device time of the build, hipEvent-timed on the context's stream */
    uint64_t clusters; /* PLOC: the live clusters summed over its rounds, the count its traffic
                          is priced by (0 for an LBVH) */
    int refitted;    /* LRH_REFIT: 1 when the tree was kept and refitted, 0 when built again */
} LrhBuildStats;

/* The message of the last failed call on this host thread. */
const char *lrh_last_error(void);

/* A context on one device with its own stream (in place of gpu_trace_create,
 * engine/gpu/trace_gpu.c:437 and gpu_create, gpu.h). stream: a hipStream_t to use, or 0 for
 * a new one of default priority. */
int lrh_create(LrhContext **out, int device, void *stream);
void lrh_destroy(LrhContext *c);
/* waits for everything queued on the context's stream (gpu_wait, engine/gpu/gpu.c:420) */
int lrh_sync(LrhContext *c);
/* the context's stream, as a hipStream_t */
void *lrh_stream(LrhContext *c);

/* Device memory for C callers that link no GPU runtime of their own (the reference's binding:
 * gpu_buffer / gpu_storage (engine/gpu/gpu_internal.h:77-83) and the readback of
 * gpu_view_trace_read (engine/gpu/view_gpu.h:117-124) do for it): a buffer of `bytes` on the context's
 * device, zeroed, ready on return; freed after the context's queued work; host <-> device
 * copies ordered after that work and waited for. */
int lrh_buffer_alloc(LrhContext *c, size_t bytes, void **d_out);
int lrh_buffer_free(LrhContext *c, void *d_buf);
int lrh_buffer_write(LrhContext *c, void *d_dst, const void *h_src, size_t bytes);
int lrh_buffer_read(LrhContext *c, void *h_dst, const void *d_src, size_t bytes);

/* triangle_bounds (engine/scene/scene.c:65-74) of n packed triangles on the device: the box
 * of q, q + u, q + v padded by 1e-3 (the boxes upload_instances makes, trace_gpu.c:1497-1502) */
int lrh_tri_boxes(LrhContext *c, const float *d_tris, int n, float *d_boxes);

/* The union of n boxes (6 floats: lo.xyz, hi.xyz) into d_bounds (device). */
int lrh_bounds(LrhContext *c, const float *d_boxes, int n, float *d_bounds);

/* One BVH over n boxes, into d_nodes[node0 .. node0 + 2n - 2] and d_aux[ref0 .. ref0 + n - 1]
 * (a leaf's shape id = its item + ref_add): build_queue + gpu_trace_build
 * (engine/gpu/trace_gpu.c:587-764) with builder LRH_LBVH (trace_build.wgsl) or LRH_PLOC
 * (trace_ploc.wgsl; n <= 2 builds an LBVH, as the reference does). h_bounds: the bounds the
 * Morton codes are taken in (6 floats, host), or NULL for the boxes' union computed on the
 * device. d_sorted (optional, 2n words): the sorted Morton codes, then the sorted items.
 * stats (optional): makes the call wait for the build and fills it. */
int lrh_build(LrhContext *c, const float *d_boxes, int n, const float *h_bounds, int builder,
              uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *d_nodes, uint32_t *d_aux,
              uint32_t *d_sorted, LrhBuildStats *stats);

/* lrh_build without allocating: no reference counterpart; for callers that cannot allocate device
 * memory (code under HIP graph capture, a serving loop's per-frame top-level rebuild, harnesses
 * that forbid hipMalloc). The same trees, bit for bit, as lrh_build. Its scratch is one fixed
 * array in the library (8 MiB on each device it runs on, held until the process ends), so n is
 * at most LRH_NOALLOC_MAX_ITEMS (LRH_EREFUSED above); it calls no allocator (stats, if given,
 * makes it wait as lrh_build does). Every call on a device shares that device's array: calls
 * that overlap on one device (two streams, or two host threads) race and must be serialised
 * by the caller; calls in order on one stream are safe. */
#define LRH_NOALLOC_MAX_ITEMS 65536
int lrh_build_noalloc(LrhContext *c, const float *d_boxes, int n, const float *h_bounds, int builder,
                      uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *d_nodes, uint32_t *d_aux,
                      uint32_t *d_sorted, LrhBuildStats *stats);

/* Closest hits (any = 0) or first hits (any = 1, shadow rays) of n rays against one binary
 * tree of triangles in the reference GPU layout: the walk of trace.wgsl traverse
 * (:270-392) for a scene of triangles. The shape id of a hit is the aux entry of its leaf.
 * Unlike the reference, the walk never drops a push: max_depth (the tree's depth, as
 * lrh_build reports it) sizes its stack, which spills from LDS to device memory. */
int lrh_trace(LrhContext *c, const uint32_t *d_nodes, uint32_t root, const uint32_t *d_aux,
              const float *d_tris, const float *d_rays, int n, int any, float *d_hits, int max_depth);

/* lrh_trace over a queue whose length is on the device (*d_count, at most max_n): the
 * extension and shadow kernels of the wavefront tracer (wf_ext, wf_shadow,
 * trace_wave.wgsl:72-84, 335-355), which the reference sizes by indirect dispatch
 * (wf_args, :320-333). */
int lrh_trace_queue(LrhContext *c, const uint32_t *d_nodes, uint32_t root, const uint32_t *d_aux, const float *d_tris,
                    const float *d_rays, int max_n, const uint32_t *d_count, int any, float *d_hits, int max_depth);

/* lrh_trace (d_count NULL) or lrh_trace_queue, also adding to d_counts[0] the node records
 * the walk fetched (the root's, an inner node's two children's, and each node popped from
 * the stack, whose two child words it reads again: 32 bytes each) and to d_counts[1] the
 * triangles it tested (48 bytes + a 4-byte aux entry each): the algorithmic traffic the
 * roofline is computed from. Slower (atomics): for measurement, not timing. */
int lrh_trace_count(LrhContext *c, const uint32_t *d_nodes, uint32_t root, const uint32_t *d_aux, const float *d_tris,
                    const float *d_rays, int max_n, const uint32_t *d_count, int any, float *d_hits, int max_depth,
                    uint64_t *d_counts);

/* Pinhole primaries through each pixel's centre of a w x h image (camera_ray,
 * engine/scene/camera.h:38-42; cam as lrs_camera_make gives it, host), row-major,
 * unit directions, t in (tmin, tmax); wf_gen's role (trace_wave.wgsl:51-70). */
int lrh_rays_primary(LrhContext *c, const float *cam, int w, int h, float tmin, float tmax, float *d_rays);

/* For each of n rays with a hit: one cosine-weighted bounce about the geometric normal
 * facing the ray (sample_cosine_hemisphere_u, engine/base/rng.h:53-70, its two numbers
 * hashed from (seed, ray index)) and one shadow ray toward the point light (host, 3 floats),
 * both from the hit point with tmin = eps, the shadow ray ending eps before the light.
 * Appended at the same slot of d_bounce and d_shadow; d_pix gets the source ray's index;
 * *d_count the number appended (queue order varies between runs; each entry does not). */
int lrh_rays_secondary(LrhContext *c, const float *d_rays, const float *d_hits, int n, const float *d_tris,
                       const float *light, float eps, uint32_t seed, float *d_bounce, float *d_shadow, uint32_t *d_pix,
                       uint32_t *d_count);

/* ---- the path tracer (engine/rt/tracer.c; the reference GPU's trace_wave.wgsl) ---- */
#define LRH_MAT_DIFFUSE 0
#define LRH_MAT_MIRROR 1
#define LRH_MAT_GLASS 2
#define LRH_MAT_LIGHT 3
#define LRH_MAT_PBR 4

/* a material (Material, engine/scene/scene.h:31-45, untextured) */
typedef struct {
    int type;
    float albedo[3], emission[3];
    float ior, metallic, roughness, transmission;
} LrhMaterial;

typedef struct LrhPtScene LrhPtScene;

/* A scene resident on the device (the role of gpu_trace_sync, engine/gpu/trace_gpu.c:1571):
 * prims (24 floats each, the 96-byte record of trace.wgsl:3-10: centre+radius | corner,
 * 0 sphere / 1 quad | edge u, material | edge v, area | unit normal, plane_d | (u x v)/|u x v|^2),
 * then triangles (lr_hip.h's layout, material in the top 6 bits); shape ids are prims first,
 * then triangles. light: the quad prim sampled as the lamp (Scene.light), or -1. eps: the
 * rays' minimum distance, margin: how far before the lamp shadow rays stop (RT_EPS,
 * RT_SHADOW_MARGIN of tracer.c:15-23, in the scene's units). The BVH over all shapes is built
 * on the device (builder LRH_LBVH or LRH_PLOC). Host arrays; the library keeps copies. */
int lrh_pt_scene(LrhContext *c, const float *prims, int nprims, const float *tris, int ntris, const LrhMaterial *mats,
                 int nmats, int light, float eps, float margin, int builder, LrhPtScene **out);
void lrh_pt_scene_free(LrhPtScene *s);
/* the scene's tree depth (what an engine's max_tree_depth must cover), its shape count, and
 * the device memory it holds (what a server's budget counts) */
int lrh_pt_scene_info(const LrhPtScene *s, int *depth, int *nshapes, size_t *bytes);

/* The sky (env.h: Scene.env): an equirect map of w x h radiance triples (w = 2h, rows top to
 * bottom, u = 0.5 looking along +z), times scale, turned by rot radians about y; rays that leave
 * the scene see it (env_eval) and next event estimation samples it by luminance x sin(theta)
 * with MIS (env_sample, env_pdf, sample_direct_env: engine/scene/env.c:54-203,
 * engine/rt/tracer.c:87-101, 415-421). rgb = NULL removes it. Beside a lamp, both are sampled at
 * every rough vertex (a shadow ray each; the sky's numbers then from the plain stream,
 * env_numbers, tracer.c:232-235). Host memory; the library keeps a copy and the sampling tables. */
int lrh_pt_scene_sky(LrhPtScene *s, const float *rgb, int w, int h, float scale, float rot);

/* Closest (any = 0) or first (any = 1) hits of n rays (d_rays, lr_hip.h's ray layout) in a
 * resident scene, through the walk its jobs use (scene_intersect / scene_occluded,
 * engine/rt/accel.h:43-45, for a batch): hit t, shape id (an instanced scene: the global
 * triangle index, 0xfffffff0 / 0xfffffff1 for a unit sphere / box), then the instance (u32;
 * other scenes: b1), b2. Asynchronous on the context's stream. */
int lrh_pt_scene_trace(LrhContext *c, const LrhPtScene *s, const float *d_rays, int n, int any, float *d_hits);

/* ---- two-level scenes (trace.wgsl:186-392: instances of meshes, unit spheres and boxes) ---- */
#define LRH_INST_SPHERE (-1)   /* the unit sphere */
#define LRH_INST_BOX (-2)      /* the box [-1, 1]^3 */
/* an instance: object -> world transform (3 x 4 rows), its kind (a mesh index, or the above),
 * its material (-1: the mesh triangles' own) */
typedef struct {
    float xform[12];
    int kind, material;
} LrhInstance;
/* A scene of instances: the meshes' triangles in object space (mesh m: triangles
 * mesh_first[m] .. mesh_first[m + 1] - 1 of mtris, lr_hip.h's layout), one BVH per mesh built
 * on the device (LRH_LBVH or LRH_PLOC), and a top-level LBVH over the instances' world boxes
 * (inverse transforms in double on the host). No lamp: light it with lrh_pt_scene_sky, and with
 * glowing shapes: instances of the unit sphere (unstretched) or box whose material glows (not
 * LRH_MAT_LIGHT, some emission), at most 16, sampled as lights at every rough vertex with MIS
 * (Scene.emitters: scene_find_emitters / scene_emitter, engine/scene/scene.c:1000-1031;
 * emit_sample / emit_pdf, engine/rt/tracer.c:120-216); their shadow rays stop 10 eps before them
 * (RT_SHADOW_MARGIN is ten RT_EPS in both of the reference's builds, tracer.c:15-23). Any
 * glowing surface adds its emission where a path finds it, weighted against that NEE (MIS). */
int lrh_pt_scene_instanced(LrhContext *c, const float *mtris, const int *mesh_first, int nmeshes, const LrhMaterial *mats,
                           int nmats, const LrhInstance *inst, int ninst, float eps, int builder, LrhPtScene **out);
/* New poses for the same instances: the top level is rebuilt on the device (per frame, as the
 * reference rebuilds its TLAS, trace_gpu.c:1371-1379), or refitted (builder LRH_REFIT). No job
 * of the scene may be running. stats (optional): the top level's build. */
int lrh_pt_scene_update(LrhPtScene *s, const LrhInstance *inst, int ninst, int builder, LrhBuildStats *stats);
/* Only the instances that moved since the last update: inst[k] is instance idx[k] (n of them,
 * each index once); their records alone are sent, in one upload scattered into place on the
 * device, and the top level is rebuilt over all the instances. The reference's
 * upload_instances sends the runs of records changed since its copy's version
 * (trace_gpu.c:1391-1452; Instance.version, scene_changed_run, scene.h:207, 359-363): its
 * caller passes those runs' indices here. n = 0: nothing is sent or rebuilt. The result equals
 * lrh_pt_scene_update with every pose. No job of the scene may be running. */
int lrh_pt_scene_move(LrhPtScene *s, const int *idx, const LrhInstance *inst, int n, int builder, LrhBuildStats *stats);
/* A dynamic mesh's new triangles (scene_mesh_changed, scene.h:385-386): ntris must be the
 * mesh's count. They are uploaded in its place, its tree is built again on the device in its
 * place (with its 4-wide copy and leaf-order triangles), and the top level over the instances
 * is rebuilt, as their world boxes follow the mesh's (upload_instances' dynamic meshes,
 * trace_gpu.c:1486-1512). The result equals a scene made with these triangles. No job of the
 * scene may be running. stats (optional): the mesh's build. */
int lrh_pt_scene_mesh(LrhPtScene *s, int mesh, const float *tris, int ntris, int builder, LrhBuildStats *stats);
/* The top level of an instanced scene to the host: its 2 ninst - 1 binary records (lr_hip.h's node
 * layout) and ninst aux entries (instance ids), and its cost (bvh_cost, engine/rt/bvh.c:464-469:
 * the records' box areas over the root's); any of the three may be NULL. */
int lrh_pt_scene_top(LrhPtScene *s, uint32_t *h_nodes, uint32_t *h_aux, double *cost);

/* Path traces spp samples (indices s0 .. s0+spp-1) of every pixel of a W x H image and adds
 * them to d_accum (float4 per pixel, row-major: the sum of the samples' rgb, and their count),
 * as gpu_frame_trace (engine/app/gpu_frame.c:49-83) and trace_image (engine/rt/tracer.c:329-356)
 * do: camera cam (lrs_camera_make's 12 floats), NEE to the lamp with MIS (mis = 1) or without,
 * max_depth bounces, samples brighter than clamp scaled down (0: none), the image's scramble
 * from seed. A sample's random numbers depend only on (seed, pixel, sample index), so any split
 * of the samples over calls gives the same sums. Returns when d_accum holds the result (the
 * samples run on the context's engine, a one-job batch). */
int lrh_pt_render(LrhContext *c, const LrhPtScene *s, const float *cam, int W, int H, int spp, uint32_t s0, uint64_t seed,
                  int max_depth, int mis, float clamp, float *d_accum);

/* The denoiser (gpu_denoise_encode, engine/gpu/denoise_gpu.c:73-99, denoise.wgsl): an
 * edge-avoiding a-trous wavelet filter steered by each pixel's variance (SVGF without its
 * temporal part) over a w x h image: d_accum float4 per pixel (the sum of its samples' rgb,
 * their count), d_moments float2 (the sums of its samples' irradiance luminance and of its
 * square), d_guide 8 floats (lrh_engine_guide's layout). passes 1..5 a-trous passes (holes
 * 1, 2, 4, 8, 16 apart; 5 is the reference's). d_out: float4 per pixel (the denoised mean, 1);
 * lamps and the sky as they are, pixels with no sample 0. Asynchronous on the context's
 * stream; device memory. */
int lrh_denoise(LrhContext *c, const float *d_accum, const float *d_moments, const float *d_guide, int w, int h,
                int passes, float *d_out);

/* The engine's BSDF on n cases (device memory): in 17 floats each (base rgb, metallic, roughness,
 * ior, transmission, outside (1/0), v xyz, l xyz, ul, u1, u2; a MAT_PBR material, the normal +z),
 * out 16 floats each: bsdf_eval's f rgb and pdf, then bsdf_pick(ul) + bsdf_sample_lobe(u1, u2):
 * ok, lobe, l xyz, weight rgb, pdf, delta, through, 0 (brdf.h:176-340). The kernel-level check of
 * the shading code, as the reference's gpu_bsdf_run (tests/test_gpu_brdf.c:21-90). Waits. */
int lrh_bsdf_cases(LrhContext *c, const float *d_in, int n, float *d_out);

/* ---- rigid bodies: the step that moves an instanced scene's dynamic instances ----
 * Spheres and boxes in a room of six walls, stepped on the device by the reference's CPU
 * solver's rules (engine/phys/rigid.h, rigid.c: contacts found once per step with speculative
 * margins, substeps of gravity, warm-started sequential impulses with Coulomb friction, a bounce
 * once from the approach speed, the relax pass). Units are the scene's (gravity 981 down: cm).
 * Sleeping and the swept test of fast bodies (rigid_ccd) are not done. One world holds at most
 * 8192 bodies. */
#define LRH_BODY_SPHERE 0
#define LRH_BODY_BOX 1
typedef struct LrhRigid LrhRigid;
/* a body (Body, rigid.h:46-64): a sphere's radius in half[0]; rot: a unit quaternion x, y, z, w
 * (body frame -> world); inv_mass 0: fixed, nothing moves it */
typedef struct {
    int shape;
    float half[3];
    float pos[3];
    float rot[4];
    float vel[3];
    float angvel[3];        /* world, radians / s */
    float inv_mass;
    float inv_inertia[3];   /* the body's own frame, diagonal */
    float restitution;      /* 0 no bounce .. 1 elastic */
    float friction;         /* Coulomb coefficient */
} LrhBody;
/* the world's settings (RigidWorld, rigid.h:80-101; rigid_init's defaults at creation: gravity
 * (0, -981, 0), no room, walls 0.4 / 0.5, 10 iterations, damping 0 / 0.1, warm start on) */
typedef struct {
    float gravity[3];
    int has_room;
    float room_lo[3], room_hi[3];
    float wall_restitution, wall_friction;
    int iterations;
    float linear_damping, angular_damping;   /* fraction of velocity lost per second */
    int warm_start;
} LrhRigidParams;
typedef struct {
    uint64_t steps;
    int contacts, pairs, colours;   /* the last step's */
    float ms;                       /* device time per step of the last lrh_rigid_steps */
    float phase_ms[4];              /* the last step's phases (device clock): pairs, contacts, colours, solver */
} LrhRigidStats;
/* rigid_init (rigid.c:122-132). max_contacts bounds a step's contacts and its listed pairs
 * (LRH_EREFUSED past them). Own device memory; steps on the context's stream. */
int lrh_rigid_create(LrhContext *c, int max_bodies, int max_contacts, LrhRigid **out);
void lrh_rigid_destroy(LrhRigid *r);
int lrh_rigid_params_get(const LrhRigid *r, LrhRigidParams *p);
int lrh_rigid_params_set(LrhRigid *r, const LrhRigidParams *p);
/* rigid_add_sphere / rigid_add_box (rigid.c:177-201): density in g / cm^3 (0: fixed); the
 * body's index, or < 0 */
int lrh_rigid_add_sphere(LrhRigid *r, const float *center, float radius, float density);
int lrh_rigid_add_box(LrhRigid *r, const float *center, const float *half, const float *rot, float density);
/* bodies [first, first + n) read back / replaced (velocities pushed, bodies fixed ...) */
int lrh_rigid_get(LrhRigid *r, int first, int n, LrhBody *out);
int lrh_rigid_set(LrhRigid *r, int first, int n, const LrhBody *in);
/* nsteps times rigid_steps(w, dt, substeps, NULL) (rigid.c:1371); waits for them */
int lrh_rigid_steps(LrhRigid *r, float dt, int substeps, int nsteps);
int lrh_rigid_stats(const LrhRigid *r, LrhRigidStats *st);

/* ---- the engine: many render jobs on one device ----
 * The role of the reference's GpuView trace loop (gpu_view_trace_start/more/passes/read,
 * engine/gpu/view_gpu.h:71-124, and gpu_frame_trace, engine/app/gpu_frame.h:14-26) for any
 * number of jobs at once. An engine holds a pool of 2^slots_log2 path slots and an image arena;
 * jobs (each a resident scene, a camera, a resolution, samples, a seed) are admitted while their
 * images fit, their paths fill the free slots as others end, and one hipGraph of several
 * wavefront iterations is replayed until the work is done. A job's image depends only on the
 * job: it is the same to the bit alone, among any others, and in the plain per-job form. */
#define LRH_EAGAIN (-5)     /* not now: no room for the job until others finish (or: timed out) */

typedef struct LrhEngine LrhEngine;

/* one render job: as gpu_frame_trace's arguments (gpu_frame.h:14-26) */
typedef struct {
    const LrhPtScene *scene;
    float cam[12];          /* lrs_camera_make's 12 floats */
    int w, h, spp;
    uint32_t s0;            /* the first sample index */
    uint64_t seed;          /* the image's scramble */
    int max_depth, mis;
    float clamp;            /* 0: none */
    int denoise;            /* 1: also the denoiser's inputs (the guide before the first sample,
                               the samples' moments), 48 more bytes a pixel; 0: none */
    float target_error;     /* > 0 (and spp > 4): stop at the first pass count K (4 <= K <= spp)
                               at which the image mean's relative standard error falls to this,
                               estimated from the first 4 passes (each pass 1 sample a pixel):
                               K = ceil(4 (e4 / target)^2), e4 that of the 4 passes' means. The
                               image is then exactly a K-spp job's. 0: all spp samples */
    float photons;          /* > 0: light tracing for caustics (light_trace_caustics, tracer.c:
                               511-665; the GPU's ltrace, trace.wgsl:1341-1511), photons x w x h
                               light paths (1 .. 16M) a pass, added to the image, the path tracer
                               then leaving their paths out (skip_caustics); only with a lamp and a
                               mirror or glass material, and no target error. The engine makes a
                               256 MB queue for light paths on its first such job. 0: none */
} LrhJob;

/* The engine's launch shape (what lrh_engine_autotune picks): work-group sizes of the walk
 * kernels (extend, shadow) and of the others (regen, shade), each 64, 128 or 256; the walk's
 * LDS stack (16, 24 or 32 entries; deeper entries spill to device memory); persistent
 * work-groups per compute unit (1..16); the queue tile a work-group claims at a time, in
 * work-group widths (1, 2, 4 or 8); wavefront iterations per captured graph (1..64); refill:
 * 0, or 1..64: the batched extension walk (single-level scenes) hands new rays to a wave's
 * finished lanes once at least that many of its 64 have finished, instead of one ray a lane
 * per tile; refill_shadow: the same for the shadow walk (single-level scenes, one light). */
typedef struct {
    int block_trace, stack, block_shade, blocks_per_cu, tile, iters_per_graph, refill, refill_shadow;
} LrhTune;

#define LRH_JOB_RUNNING 1
#define LRH_JOB_DONE 2

typedef struct {
    int state;              /* LRH_JOB_RUNNING or LRH_JOB_DONE */
    int w, h, spp;
    uint64_t paths;         /* w * h * spp */
    double t_submit_ms;     /* when it was admitted (lrh_now_ms's clock) */
    double t_done_ms;       /* when its last sample ended, from the device's clock (0 while running) */
    int spp_done;           /* the samples per pixel it takes: spp, or K with a target error (0 while
                               its first 4 passes decide) */
    float rel_error;        /* with a target error: the image mean's relative standard error at
                               spp_done passes, as estimated (e4 sqrt(4 / K)); else 0 */
} LrhJobInfo;

typedef struct {
    uint64_t rays_ext;      /* closest-hit rays traced (every queue entry walked) */
    uint64_t rays_shadow;   /* any-hit shadow rays traced */
    uint64_t iterations;    /* wavefront iterations run */
    uint64_t jobs_done;
    uint64_t node_fetches;  /* with counting on: node records read by the extension walks */
    uint64_t shape_tests;   /* and shapes tested */
    double kernel_ms[4];    /* with profiling on: device time of regen, extend, shade, shadow */
    uint64_t kernel_launches[4];
    /* with counting on: the extension walks' wave steps (node steps, shape steps; a step is one
     * wave's pass through that part of the walk, whatever lanes take part), so that
     * node_fetches / (64 wave_node_steps) is the share of lanes at work in a node step */
    uint64_t wave_node_steps;
    uint64_t wave_shape_steps;
    /* and the shadow walks' node records, shapes and wave steps */
    uint64_t shadow_node_fetches;
    uint64_t shadow_shape_tests;
    uint64_t shadow_wave_node_steps;
    uint64_t shadow_wave_shape_steps;
    /* and the node steps in which every lane at work is at one node of one tree (single-level
     * walks), for the extension and the shadow walks */
    uint64_t uniform_node_steps;
    uint64_t shadow_uniform_node_steps;
    /* light tracing for caustics (LrhJob.photons): the light paths traced, their closest-hit rays
     * and their rays towards the camera, with counting on the node records and shapes their
     * walks read, its launches (one a job), and with profiling on its device time */
    uint64_t light_paths;
    uint64_t light_rays;
    uint64_t light_shadow_rays;
    uint64_t light_node_fetches;
    uint64_t light_shape_tests;
    uint64_t light_launches;
    double light_ms;
} LrhEngineStats;

/* host clock (steady, ms) that job times are given in (clock_ms, engine/base/clock.h:6) */
double lrh_now_ms(void);

/* slots_log2 in 10..22; max_tree_depth: the deepest tree any job's scene may have (its
 * walk stack); image_bytes: the arena for job images (24 bytes a pixel). Own stream. */
int lrh_engine_create(LrhContext *c, int slots_log2, int max_tree_depth, size_t image_bytes, LrhEngine **out);
void lrh_engine_destroy(LrhEngine *e);
int lrh_engine_tune_set(LrhEngine *e, const LrhTune *t);
int lrh_engine_tune_get(const LrhEngine *e, LrhTune *t);
/* counting on: the extension and shadow walks also count node records, shapes and their waves'
 * steps (slower kernels) */
int lrh_engine_counting(LrhEngine *e, int on);
/* profiling on: lrh_engine_pump launches the iterations kernel by kernel between hipEvents on
 * the engine's stream (no graph) and sums each kernel's time into the stats */
int lrh_engine_profiling(LrhEngine *e, int on);
/* Admits a job (its image cleared): LRH_EAGAIN if its image or a job slot cannot be had until
 * others are released. *id names it. Its paths are issued to the device in chunks of at most
 * one pool of slots, always from the admitted job with the fewest paths left to issue
 * (shortest remaining first), just ahead of what the device consumes. */
int lrh_engine_submit(LrhEngine *e, const LrhJob *job, int64_t *id);
/* Runs up to `graphs` captured graphs of the engine's iterations (fewer if the work ends) and
 * takes in the progress they report. Returns the number of jobs still running (>= 0). */
int lrh_engine_pump(LrhEngine *e, int graphs);
/* pump until every admitted job is done */
int lrh_engine_drain(LrhEngine *e);
/* The reference's shape (gpu_frame_trace over one view, engine/app/gpu_frame.c:49-83): the
 * job alone, one pass of one sample per pixel after another, each pass a grid over its paths,
 * the host waiting for the job's end. The engine must be idle. */
int lrh_engine_plain(LrhEngine *e, const LrhJob *job, int64_t *id);
int lrh_engine_job(LrhEngine *e, int64_t id, LrhJobInfo *info);
/* the job's image into d_out: float4 per pixel (the sum of its samples' rgb, their count),
 * stored (add = 0) or added (add = 1); returns when d_out holds it */
int lrh_engine_image(LrhEngine *e, int64_t id, float *d_out, int add);
/* A finished job submitted with denoise = 1: its guide into d_guide (8 floats a pixel: the
 * normal facing the camera of the first diffuse surface seen through the pixel's centre, through
 * mirrors and glass, and its distance along the way | the albedo of the way to it, the kind: 0
 * a surface, 1 a lamp, a bright glow or the sky; trace.wgsl:1034-1107) and its moments into
 * d_moments (2 floats a pixel: the sums of its samples' irradiance luminance, the sample over
 * the guide's albedo on a surface, and of its square; trace.wgsl:1301-1308). Waits. */
int lrh_engine_guide(LrhEngine *e, int64_t id, float *d_guide, float *d_moments);
/* A finished job submitted with denoise = 1, denoised (lrh_denoise on its image, moments and
 * guide) into d_out: float4 per pixel (the denoised mean rgb, 1). Waits. */
int lrh_engine_denoise(LrhEngine *e, int64_t id, int passes, float *d_out);
/* forgets a finished job: its image memory and slot go back to the engine */
int lrh_engine_release(LrhEngine *e, int64_t id);
int lrh_engine_stats(LrhEngine *e, LrhEngineStats *st);
/* Times the sample jobs (their spp as given: keep it small) under candidate launch shapes,
 * one choice at a time (walk work-group size x stack, then work-groups per CU, then the tile,
 * then the shading work-group size), keeps the fastest in rays per second and sets it. */
int lrh_engine_autotune(LrhEngine *e, const LrhJob *sample, int n, LrhTune *best, double *mrays);
/* lrh_pt_render of one job with all its options: light-traced caustics (photons > 0, as
 * gpu_frame_trace's photons argument, engine/app/gpu_frame.c:49-52, passed on to the GPU's
 * caustics pass, engine/gpu/trace_gpu.c:1587-1594), a target error. Its sums (and sample
 * counts) added to d_accum (float4 per pixel); returns when d_accum holds them. Not with
 * denoise (its outputs come from an engine: lrh_engine_denoise). */
int lrh_pt_render_job(LrhContext *c, const LrhJob *job, float *d_accum);

/* ---- the photon cache: the viewer's caustics between frames ----
 * gpu_trace_photons / gpu_trace_photons_clear / gpu_trace_photon_paths (engine/gpu/trace_gpu.c:
 * 1964-1996) with the light tracer's cache mode (deposit, pc_claim: trace.wgsl:1359-1404), its
 * fading (trace_decay.wgsl) and its lookup (trace_resolve.wgsl:43-80). A world-space hash grid of
 * 2^20 cells (32 MB): cells of three sizes (cell, x4, x16), one set per main axis of the
 * surface's normal and its sign; each holds the power the light paths left there (integer sums:
 * the same table whatever the order) and its photons. Light paths start at the scene's lamp as
 * the light tracer's do and leave their power at every rough vertex after a mirror or glass. */
typedef struct LrhPhotons LrhPhotons;
/* cell: the finest cells' size (the reference's T.lt.z) */
int lrh_photons_create(LrhContext *c, float cell, LrhPhotons **out);
void lrh_photons_destroy(LrhPhotons *pc);
int lrh_photons_clear(LrhPhotons *pc);
/* npaths light paths (path i: PCG32 stream i, sample index 2^31 + s0) of a one-level scene,
 * max_depth vertices each, into the cache (asynchronous, on the context's stream) */
int lrh_photons_trace(LrhPhotons *pc, const LrhPtScene *s, uint32_t npaths, uint32_t s0, uint64_t seed, int max_depth);
/* every cell's power and photons times k in [0, 1] (the light paths it holds too); a cell whose
 * photons faded to nothing gives its slot back when the next slot is free */
int lrh_photons_fade(LrhPhotons *pc, float k);
/* the light paths the cache holds (their number, faded with it) */
double lrh_photons_paths(const LrhPhotons *pc);
/* each pixel's centre ray (camera cam); where it first meets a rough surface, the caustics'
 * radiance there from the cache: albedo / pi x the power of the finest cells around the point
 * (bilinearly) holding at least 16 photons, over their area and the light paths; d_out: float4
 * per pixel (rgb, 1; 0 where no rough surface is seen). Returns when d_out holds it. */
int lrh_photons_resolve(LrhPhotons *pc, const LrhPtScene *s, const float *cam, int W, int H, float *d_out);
/* the table to the host: 2^20 slots of 4 x 64 bits (rgb sums in 2^-8 units; the photons in 1/16
 * units in the low 32 bits of the fourth word, the cell's tag in its high 32 bits, 0 free) */
int lrh_photons_read(LrhPhotons *pc, void *h_table);

/* ---- serving: a request loop over an engine ----
 * Scenes registered with the server stay resident; their bytes and the images of admitted jobs
 * count against a budget of device memory: the image arena is made in chunks as admitted images
 * need it, so resident scenes plus the arena never hold more than the budget (the engine's path
 * state, fixed by slots_log2 at ~360 bytes a slot, a finished image's staging buffer and, once a
 * job traces light paths, their 256 MB queue come on top of it). Requests (thread-safe, from any host thread) wait
 * in order until their image fits, then join the running batch at once: their paths take free
 * slots as others end (continuous batching). A worker thread drives the engine. */
typedef struct LrhServer LrhServer;

typedef struct {
    int64_t submitted, admitted, done, waiting;
    uint64_t rays;              /* rays traced (closest-hit + shadow) */
    double busy_ms;             /* time with jobs on the device */
    double p50_ms, p99_ms;      /* latency (submit -> last sample) of the jobs done */
    size_t budget, in_use;      /* bytes: the budget; device memory held by the resident scenes and the image arena */
} LrhServerStats;

int lrh_server_create(LrhContext *c, int slots_log2, int max_tree_depth, size_t budget_bytes, const LrhTune *tune,
                      LrhServer **out);
/* a resident scene (LRH_EREFUSED if it does not fit the budget) */
int lrh_server_add_scene(LrhServer *s, const LrhPtScene *scene);
int lrh_server_submit(LrhServer *s, const LrhJob *job, int64_t *id);
/* waits up to timeout_ms (< 0: no limit) for the job: 0 done, LRH_EAGAIN not yet */
int lrh_server_wait(LrhServer *s, int64_t id, double timeout_ms, LrhJobInfo *info);
/* a finished job's image into host memory h_out (float4 per pixel: sum, count) */
int lrh_server_image(LrhServer *s, int64_t id, float *h_out);
int lrh_server_release(LrhServer *s, int64_t id);
int lrh_server_stats(LrhServer *s, LrhServerStats *st);
void lrh_server_destroy(LrhServer *s);

#ifdef __cplusplus
}
#endif
#endif
