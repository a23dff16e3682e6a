/* lr_scene.h - the host side of the C ABI: seeded synthetic scenes in the device layouts of
 * lr_hip.h, and the pinhole camera. Plain C, host memory only. */
#ifndef LR_SCENE_H
#define LR_SCENE_H

#include <stdint.h>

#include "lr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A triangle soup for BVH stress: n triangles whose centroids are uniform in the unit cube,
 * each with two edges of log-normal length (mu = ln 0.002, sigma = 0.5) along uniformly
 * random directions, the corners placed so that the centroid is the drawn point. Drawn from
 * PCG32 (engine/base/rng.h:13-28) seeded (seed, stream 1), in double, stored as f32 in the
 * 48-byte triangle layout; material m = the triangle's index mod nmat (0 if nmat <= 1),
 * normals flat. */
void lrs_gen_soup(int n, uint64_t seed, int nmat, float *tris);

/* camera_make (engine/scene/camera.h:21-36): 12 floats, origin | upper_left | horizontal |
 * vertical, each 3 floats; computed in double, stored as f32 */
void lrs_camera_make(const double from[3], const double at[3], const double up[3], double vfov_deg, double aspect,
                     float cam[12]);

/* The sampler's 64 x 64 blue-noise ranks (void-and-cluster, engine/base/sampler.c:86-132, as
 * the reference's float build makes them) */
void lrs_blue_noise(uint16_t *rank);
/* The BSDF's tables (E[32][32], B[32][32], E_avg[32], B_avg[32]; engine/scene/brdf_tables.h's
 * layout), computed by the quadrature tools/gen_brdf_tables.py documents */
void lrs_brdf_tables(float *out);
/* make_quad (engine/scene/scene.c:26-40) into the 24-float prim record of lrh_pt_scene */
void lrs_make_quad(const float q[3], const float u[3], const float v[3], int material, float prim[24]);

/* The Cornell-style box of the path-tracing workload, in a unit box (the reference's 555-unit
 * box, engine/scene/scene.c:438-464, scaled by 1/555): 5 walls as 10 triangles, two boxes
 * turned +15 and -18 degrees about y (24 triangles), each of the 17 flat surfaces Lambertian
 * with an albedo uniform in [0.2, 0.8] per channel (PCG32 seeded (seed, 1)); a 0.25 x 0.25 quad
 * lamp of radiance 15 just under the ceiling, facing down. Fills tris (34 x 12 floats), prims
 * (1 x 24: the lamp), mats (18: the surfaces, then the lamp's), the camera of the reference's
 * default view ((278, 278, -800) toward (278, 278, 0), vfov 40, scaled) for a w x h image. */
void lrs_cornell(uint64_t seed, int w, int h, float *tris, float *prims, LrhMaterial *mats, float cam[12]);

/* ---- the batched (many jobs in one launch) and serving workloads ---- */
#define LRS_BATCH_SOUP_MATS 14

/* A scene of the batched workload: the soup of lrs_gen_soup (n triangles, seed) over 14 GGX
 * materials drawn from PCG32 (seed, stream 3): base colour uniform in [0.2, 0.9] per channel,
 * roughness uniform in [0.05, 1], metallic with probability 0.4, and two of them (every 7th)
 * clear glass of IOR 1.5; a floor quad (2 triangles, material 14, diffuse 0.5) at y = -0.05 over
 * [-2, 3]^2; a 0.5 x 0.5 quad lamp of radiance 15 (material 15) at y = 1.6 over the cube's
 * centre, facing down. tris: (n + 2) x 12 floats; prims: 1 x 24 (the lamp, prim 0); mats: 16. */
void lrs_batch_scene(int n, uint64_t seed, float *tris, float *prims, LrhMaterial *mats);
/* a job's camera: from a direction drawn from PCG32 (seed, stream 4) with y >= 0.15, 2.2 from
 * the soup's centre (0.5, 0.5, 0.5), looking at it; vfov 40, aspect w / h */
void lrs_batch_camera(uint64_t seed, int w, int h, float cam[12]);
/* The job mix, from PCG32 (seed, stream 5): per job a triangle count log-uniform in [1000,
 * 500000], a resolution uniform over 256^2, 512^2, 1280x720, 1920x1080, and spp uniform over
 * {1, 4, 16, 64} (spp_lo = 4: over {4, 16, 64}, the serving mix). */
void lrs_batch_plan(int njobs, uint64_t seed, int spp_lo, int *ntris, int *w, int *h, int *spp);

/* ---- the sky (engine/scene/env.c) ---- */
/* env_procedural_sky (env.c:510-537) with the sun's radiance a parameter: a w x w/2 equirect
 * map (rows top to bottom, u = 0.5 looking along +z), blue above with haze at the horizon, a
 * glow and a 1-degree sun at (sun_elev, sun_azim) degrees, a grey ground below; f32 as the
 * reference's float build computes it. sun_le = (2600, 2400, 2100) is the reference's. */
void lrs_sky(int w, float sun_elev, float sun_azim, const float sun_le[3], float *rgb);
/* The sky's sampling tables (make_weights and prepare_base, env.c:125-164, 335-364): marg
 * (h + 1), cond (h x (w + 1)), prob (w x h: each pixel's probability, pixel_prob), and the
 * weights' sum. 0, or -1 on bad arguments or no memory. */
int lrs_sky_tables(const float *rgb, int w, int h, float *marg, float *cond, float *prob, double *sum);

/* ---- meshes for instancing (object space; each returns its triangle count) ---- */
/* a unit UV sphere: 2 x slices x (stacks - 1) triangles facing out (slices 30, stacks 17: 960) */
int lrs_uv_sphere(int slices, int stacks, int material, float *tris);
/* the box [-1, 1]^3: 12 triangles facing out */
int lrs_box_mesh(int material, float *tris);
/* the convex hull of npts points drawn on the unit sphere from PCG32 (seed, stream 6): 2 npts - 4
 * triangles facing out (64 points: 124); -1 if out of memory */
int lrs_hull(int npts, uint64_t seed, int material, float *tris);
/* a seeded blob: a UV sphere of about ntris triangles (room for ntris + 4 * sqrt(ntris) + 64),
 * its radius bumped by four waves drawn from PCG32 (seed, stream 7) */
int lrs_blob(int ntris, uint64_t seed, int material, float *tris);

/* the inverse of an affine 3 x 4 transform (rows), in double, stored as f32: -1 if singular */
int lrs_affine_inverse(const float fwd[12], float inv[12]);

#ifdef __cplusplus
}
#endif
#endif
