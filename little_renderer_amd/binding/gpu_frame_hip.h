/* gpu_frame_hip.h - what gpu_frame_hip.c adds to engine/app/gpu_frame.h on liblr_hip.so: batches
 * of the scene-wide ray queries of engine/rt/accel.h:43-45 on the device, and the samples per
 * pixel of the view's last image (the role of gpu_view_trace_passes, engine/gpu/view_gpu.h:113,
 * which a build without the WebGPU layer keeps as a stub). Compiled against the reference's
 * headers (it uses Scene, Ray, Hit, GpuView). */
#ifndef LR_GPU_FRAME_HIP_H
#define LR_GPU_FRAME_HIP_H

#include "app/gpu_frame.h"
#include "rt/intersect.h"

/* scene_intersect for n rays at once, through the view's device copy of s (made or brought up
 * to date as gpu_frame_trace does): found[i] = 1 and hits[i] filled as set_hit fills it
 * (t, p, the geometric normal facing the ray as n and ng, front_face, prim, inst, uv), or
 * found[i] = 0. 0 on success, 1 if the device cannot hold the scene or failed (the caller
 * asks the CPU's scene_intersect). */
int gpu_scene_intersect(GpuView *view, const Scene *s, const Ray *rays, int n, real tmin, real tmax, Hit *hits,
                        int *found);
/* scene_occluded for n rays at once: occluded[i] = 1 if anything is hit in (tmin, tmax) */
int gpu_scene_occluded(GpuView *view, const Scene *s, const Ray *rays, int n, real tmin, real tmax, int *occluded);
/* the samples per pixel of the view's last gpu_frame_trace image */
int gpu_frame_passes(const GpuView *view);

#endif
