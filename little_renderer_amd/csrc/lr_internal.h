// lr_internal.h - the library's host-side context and the launchers each .hip file provides.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/lr_hip.h"

// parameters of one build, on the device: the bounds' low corner, 1 / their extent (as
// build_queue computes it, trace_gpu.c:603-611), and the high corner
struct LrBuildParamsDev {
    float lo[4];
    float ext[4];
    float hi[4];
};

// the scratch memory of one context, grown on demand and kept between calls
struct LrScratch {
    void *base = nullptr;
    size_t bytes = 0;
};

enum { LR_SCRATCH_BUILD = 0, LR_SCRATCH_TRACE = 1, LR_SCRATCH_PT = 2, LR_SCRATCH_COUNT = 3 };

struct LrhContext {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    LrScratch scratch[LR_SCRATCH_COUNT];
    uint32_t *flags = nullptr;   // device words: [0] a walk's stack overflowed (never expected)
    hipEvent_t ev[2] = {nullptr, nullptr};
    int ploc_small = 1024;   // PLOC: the clusters from which the rounds run in one work-group
    float last_build_ms = 0.0f;
    LrhEngine *engine = nullptr;  // lrh_pt_render's engine, made on first use
};

// error reporting: a thread-local message, negative codes (lr_hip.h)
int lr_fail(int code, const char *fmt, ...);
#define LR_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return lr_fail(LRH_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// scratch `which` of at least `bytes` (256-byte aligned pieces are carved by the callers);
// growing it waits for the stream first, so no kernel still uses the old memory
int lr_scratch_reserve(LrhContext *c, int which, size_t bytes);

static inline size_t lr_align(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- lr_build.hip ----
int lr_launch_tri_boxes(LrhContext *c, const float *tris, int n, float *boxes);
int lr_launch_bounds(LrhContext *c, const float *boxes, int n, LrBuildParamsDev *out, void *tmp);
size_t lr_bounds_tmp_bytes();
// the build's scratch for n items (both builders)
size_t lr_build_scratch_bytes(int n);
int lr_build_run(LrhContext *c, const float *boxes, int n, const float *host_bounds, int builder,
                 uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux,
                 uint32_t *sorted_codes_out, uint32_t *sorted_items_out, LrhBuildStats *stats);

// the build with its scratch in the library's own device array (lrh_build_noalloc): the array's
// device address, taken once per device with that device current (from lrh_create)
int lr_noalloc_init(int device);
int lr_build_noalloc(LrhContext *c, const float *boxes, int n, const float *host_bounds, int builder, uint32_t node0,
                     uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux, uint32_t *sorted,
                     LrhBuildStats *stats);

// ---- lr_denoise.hip ----
size_t lr_denoise_scratch_bytes(int w, int h);
// the denoiser on stream s: accum (float4 sum, count), moments (float2), guide (2 float4 a
// pixel) -> out (float4 rgb, 1); scratch: lr_denoise_scratch_bytes
int lr_denoise_launch(hipStream_t s, const float *accum, const float *moments, const float *guide, int w, int h, int passes,
                      void *scratch, float *out);

// ---- lr_trace.hip ----
int lr_launch_trace(LrhContext *c, const uint32_t *nodes, uint32_t root, const uint32_t *aux, const float *tris,
                    const float *rays, int nrays, const uint32_t *count, int any, float *hits, int max_depth,
                    unsigned long long *counts = nullptr);

// ---- lr_rays.hip ----
int lr_launch_rays_primary(LrhContext *c, const float *cam, int w, int h, float tmin, float tmax, float *rays);
int lr_launch_clear(LrhContext *c, uint32_t *p, int n);
int lr_launch_rays_secondary(LrhContext *c, const float *rays, const float *hits, int n, const float *tris,
                             const float *light, float eps, uint32_t seed, float *bounce, float *shadow, uint32_t *pix,
                             uint32_t *count);
