/* lro_pt.h - the oracle's path tracer (lro_pt.c). TEST INFRASTRUCTURE ONLY (see lro.h). */
#ifndef LRO_PT_H
#define LRO_PT_H

#include <stdint.h>

#include "lro.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LRO_SAMPLER_PAIRS 13                 /* sampler.h:29-33 */
#define LRO_SAMPLE_LIGHT(d) (1 + 3 * (d))
#define LRO_SAMPLE_DIR(d) (2 + 3 * (d))
#define LRO_SAMPLE_MISC(d) (3 + 3 * (d))
#define LRO_BRDF_N 32

enum { LRO_MAT_DIFFUSE, LRO_MAT_MIRROR, LRO_MAT_GLASS, LRO_MAT_LIGHT, LRO_MAT_PBR };   /* scene.h:23-29 */

/* a material (scene.h:31-45, without textures) */
typedef struct {
    int type;
    float albedo[3], emission[3];
    float ior, metallic, roughness, transmission;
} LroMat;

/* a primitive in the GPU's 96-byte record (trace.wgsl:3-10): c = centre, radius; q = corner,
 * w 0 sphere / 1 quad; u = edge, w material; v = edge, w area; n = unit normal, w plane_d;
 * w = (u x v) / |u x v|^2 */
typedef struct {
    float c[4], q[4], u[4], v[4], n[4], w[4];
} LroPrim;

/* the sky (engine/scene/env.h): an equirect map w x h (w = 2h) and its sampling tables */
typedef struct {
    int w, h;
    const float *rgb;                     /* w * h * 3, rows top to bottom */
    float *marg, *cond, *prob;            /* h + 1; h * (w + 1); w * h (pixel_prob) */
    double sum;
    float scale, rot, rot_c, rot_s;
} LroEnv;
/* the tables of env_prepare's sampling part (env.c:125-164, 335-364); 0 or -1 */
int lro_env_prepare(LroEnv *e, const float *rgb, int w, int h, float scale, float rot);
void lro_env_free(LroEnv *e);
/* env_eval, env_sample, env_pdf (env.c:54-203) */
void lro_env_eval(const LroEnv *e, const float d[3], float out[3]);
void lro_env_sample(const LroEnv *e, float u1, float u2, float d[3], float *pdf, float le[3]);
float lro_env_pdf(const LroEnv *e, const float d[3]);

/* an instance of a two-level scene: world -> object rows (3 x 4), its kind (a mesh index >= 0,
 * LRO_INST_SPHERE: the unit sphere, LRO_INST_BOX: the box [-1, 1]^3), its material (-1: the
 * mesh triangles' own) */
#define LRO_INST_SPHERE (-1)
#define LRO_INST_BOX (-2)
typedef struct {
    float inv[12];
    int kind, material;
} LroInst;

/* a glowing shape sampled as a light (Emitter, engine/scene/scene.h:369; scene_emitter,
 * scene.c:1015-1031): a ball (centre c, radius r) or a box (centre c, half axes a[0..2] as the
 * columns of its transform), its area and radiance */
#define LRO_MAX_EMITTERS 16
typedef struct {
    int ball;
    float c[3], a[9], r, area, le[3];
} LroEmitter;

typedef struct {
    const LroPrim *prims; int nprims;
    const float *tris; int ntris;        /* 12 floats each, material in the top 6 bits of word 11 */
    const LroMat *mats; int nmats;
    int light;                            /* the lamp: a quad prim, or -1 */
    float eps, margin;                    /* ray offset and shadow margin (tracer.c:15-23) */
    const uint16_t *mask;                 /* the blue-noise ranks (64 x 64) */
    const float *tab;                     /* the BSDF tables: E[32][32], B[32][32], E_avg[32], B_avg[32] */
    const LroEnv *env;                    /* the sky, or NULL (black) */
    /* instances (a two-level scene): the meshes' triangles in object space (mesh m: triangles
       mesh_first[m] .. mesh_first[m + 1] - 1 of mtris) and the instances, besides the shapes above */
    const float *mtris; const int *mesh_first; int nmeshes;
    const LroInst *inst; int ninst;
    /* optional: each mesh's box in object space (lo.xyz, hi.xyz), padded: an instance whose box
       the ray misses is skipped (a cull that drops no hit: the brute force stays exact) */
    const float *mesh_box;
    /* optional: a binary tree over tris (lro.h's node layout, aux = triangle index), walked for
       the triangles instead of the brute force (nodes NULL: brute force) */
    const uint32_t *tnodes, *taux;
    /* the glowing shapes (Scene.emitters) and each instance's slot among them (-1: none) */
    const LroEmitter *emit; int nemit;
    const int *inst_emit;
    /* light tracing on (tracer.h:20): the path tracer leaves out the lamp reached right after a
       mirror or glass on a path whose first vertex was rough (light tracing adds those) */
    int skip_caustics;
} LroScene;

/* scene_find_emitters + scene_emitter (engine/scene/scene.c:1000-1031): the instances that are
 * glowing shapes (a unit sphere whose three axes have one length, or a unit box, whose material
 * glows: not a lamp, some emission), at most LRO_MAX_EMITTERS, from each instance's object ->
 * world rows (fwd, 12 floats), kind and material; slot[i] = its emitter or -1. Returns how many. */
int lro_find_emitters(const float *fwd, const int *kind, const int *material, int ninst, const LroMat *mats,
                      int nmats, LroEmitter *out, int *slot);

/* The closest hit of n rays (8 floats each) in a scene, by brute force over every shape and
 * every instance's shapes: hit = t, shape id (a world shape, or the instance's local triangle,
 * 0xfffffff0 its unit sphere, 0xfffffff1 its unit box; 0xffffffff: no hit), instance (u32,
 * 0xffffffff: none), unused */
void lro_scene_trace(const LroScene *s, const float *rays, int n, int any, float *hits);

uint32_t lro_sampler_hash(uint32_t x);
uint32_t lro_sampler_owen(uint32_t x, uint32_t seed);
uint32_t lro_sampler_sobol1(uint32_t i);
void lro_blue_noise(uint16_t *rank);
void lro_sampler_2d(const uint16_t *mask, int x, int y, uint32_t index, uint32_t seed, int dim, float out[2]);
/* one BSDF case: eval (f.rgb, pdf) then pick(ul) + sample_lobe(u1, u2): ok, lobe, l.xyz, w.rgb,
 * pdf, delta, through */
void lro_bsdf_case(const LroMat *m, int outside, const float *tab, const float n[3], const float v[3], const float l[3],
                   float ul, float u1, float u2, float out[16]);
/* lro_bsdf_case on n cases of 17 floats (material_pbr: base rgb, metallic, roughness, ior,
 * transmission; outside; v; l; ul, u1, u2; the normal +z) into 16 floats each */
void lro_bsdf_cases(const float *in, int n, const float *tab, float *out);
/* the per-(pixel, sample) PCG32 stream */
void lro_pt_seed(LroRng *rng, uint64_t pixel, uint32_t k, uint64_t seed);
/* adds samples s0 .. s0+ns-1 of every pixel (rgb sums) into sum */
void lro_pt_image(const LroScene *s, const float cam[12], int W, int H, uint32_t s0, int ns, uint64_t seed, float clamp,
                  int max_depth, int mis, float *sum);
/* light tracing for caustics: npass passes of nlp light paths, their splats (fixed point, 2^-28)
 * added to acc (H x W x 3 int64) */
void lro_light_trace(const LroScene *s, const float cam[12], int W, int H, uint32_t nlp, uint32_t s0, int npass,
                     uint64_t seed, int max_depth, long long *acc);
/* the photon cache (lrh_photons_*): a table of LRO_PC_SLOTS slots of 4 x 64 bits */
#define LRO_PC_SLOTS (1u << 20)
uint64_t lro_pc_key(int x, int y, int z, int axis, int level);
void lro_pc_deposit(uint64_t *table, float cell, const float p[3], const float n[3], const float c[3]);
void lro_pc_deposit_many(uint64_t *table, float cell, int count, const float *p, const float *n, const float *c);
void lro_pc_fade(uint64_t *table, double k);
void lro_photons_trace(const LroScene *s, uint64_t *table, float cell, uint32_t npaths, uint32_t s0, uint64_t seed,
                       int max_depth);
void lro_pc_resolve(const LroScene *s, const uint64_t *table, float cell, double paths, const float cam[12], int W, int H,
                    float *out);
/* the denoiser's guide of a W x H image (8 floats a pixel: n.xyz, z | albedo.rgb, kind) */
void lro_guide(const LroScene *s, const float cam[12], int W, int H, float *out);
/* lro_pt_image, and each pixel's moments of its samples' irradiance luminance (2 floats) */
void lro_pt_image_moments(const LroScene *s, const float cam[12], int W, int H, uint32_t s0, int ns, uint64_t seed,
                          float clamp, int max_depth, int mis, const float *guide, float *sum, float *moments);
/* the denoiser (lrh_denoise's filter): accum float4, moments float2, guide 8 floats -> out float4 */
void lro_denoise(const float *accum, const float *moments, const float *guide, int W, int H, int passes, float *out);
/* its packing: IEEE-754 halves (round to nearest even), the normal on an octahedron in two snorm16 */
uint16_t lro_f2h(float f);
float lro_h2f(uint16_t h);
void lro_oct_pack(const float n[3], uint16_t out[2]);
void lro_oct_unpack(const uint16_t in[2], float n[3]);
/* the same for the pixels x0 .. x0+cw-1, y0 .. y0+ch-1 of the W x H image only (sum: ch x cw) */
void lro_pt_window(const LroScene *s, const float cam[12], int W, int H, int x0, int y0, int cw, int ch, uint32_t s0,
                   int ns, uint64_t seed, float clamp, int max_depth, int mis, float *sum);

#ifdef __cplusplus
}
#endif
#endif
