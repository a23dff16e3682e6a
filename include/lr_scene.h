/* lr_scene.h - the host side of the C ABI: seeded synthetic scenes in the device layouts of
 * lr_hip.h, and the pinhole camera. Plain C, host memory only. */
#ifndef LR_SCENE_H
#define LR_SCENE_H

#include <stdint.h>

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

#ifdef __cplusplus
}
#endif
#endif
