// lr_api.hip - the C ABI (include/lr_hip.h): contexts, errors, scratch memory, and the entry
// points that check their arguments and hand over to the launchers.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "lr_internal.h"

static thread_local char g_err[512];

int lr_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int lr_scratch_reserve(LrhContext *c, int which, size_t bytes)
{
    LrScratch &s = c->scratch[which];
    if (s.base && s.bytes >= bytes) return 0;
    if (s.base) {
        LR_HIP(hipStreamSynchronize(c->stream));
        LR_HIP(hipFree(s.base));
        s.base = nullptr;
        s.bytes = 0;
    }
    const size_t want = lr_align(bytes + bytes / 8 + 4096);
    if (hipMalloc(&s.base, want) != hipSuccess) {
        (void)hipGetLastError();
        s.base = nullptr;
        return lr_fail(LRH_ENOMEM, "scratch: %zu bytes of device memory could not be had", want);
    }
    // the build's counters start at zero (its kernels leave them so). Cleared on the
    // context's stream, so that the clear is ordered before the kernels that use the memory
    // (a plain hipMemset runs on the null stream, which a non-blocking stream does not wait for)
    LR_HIP(hipMemsetAsync(s.base, 0, want, c->stream));
    LR_HIP(hipStreamSynchronize(c->stream));
    s.bytes = want;
    return 0;
}

extern "C" {

const char *lrh_last_error(void) { return g_err; }

int lrh_create(LrhContext **out, int device, void *stream)
{
    if (!out) return lr_fail(LRH_EINVAL, "lrh_create: out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return lr_fail(LRH_EDEVICE, "lrh_create: no HIP device");
    }
    if (device < 0 || device >= count) return lr_fail(LRH_EINVAL, "lrh_create: device %d of %d", device, count);
    LR_HIP(hipSetDevice(device));
    LrhContext *c = new LrhContext();
    c->device = device;
    if (stream) c->stream = (hipStream_t)stream;
    else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return lr_fail(LRH_EDEVICE, "lrh_create: no stream");
        }
        c->own_stream = true;
    }
    if (hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess ||
        hipMalloc(&c->flags, 256) != hipSuccess || hipMemsetAsync(c->flags, 0, 256, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        lrh_destroy(c);
        return lr_fail(LRH_EDEVICE, "lrh_create: events or flags could not be made");
    }
    // the device address of the no-allocation build's scratch, once per device (this one current)
    const int rc = lr_noalloc_init(device);
    if (rc) {
        lrh_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

void lrh_destroy(LrhContext *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    lrh_engine_destroy(c->engine);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &s : c->scratch)
        if (s.base) (void)hipFree(s.base);
    if (c->flags) (void)hipFree(c->flags);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void *lrh_stream(LrhContext *c) { return c ? (void *)c->stream : nullptr; }

int lrh_sync(LrhContext *c)
{
    if (!c) return lr_fail(LRH_EINVAL, "lrh_sync: no context");
    LR_HIP(hipStreamSynchronize(c->stream));
    uint32_t f = 0;
    LR_HIP(hipMemcpy(&f, c->flags, 4, hipMemcpyDeviceToHost));
    if (f) {
        LR_HIP(hipMemsetAsync(c->flags, 0, 4, c->stream));
        LR_HIP(hipStreamSynchronize(c->stream));
        return lr_fail(LRH_EREFUSED, "a BVH walk needed more stack than the tree's depth allows (depth too small?)");
    }
    return 0;
}

int lrh_buffer_alloc(LrhContext *c, size_t bytes, void **d_out)
{
    if (!c || !d_out || bytes == 0) return lr_fail(LRH_EINVAL, "lrh_buffer_alloc: bad arguments");
    *d_out = nullptr;
    LR_HIP(hipSetDevice(c->device));
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return lr_fail(LRH_ENOMEM, "lrh_buffer_alloc: %zu bytes of device memory could not be had", bytes);
    }
    // zeroed on the context's stream, and waited for: the buffer is ready when the call returns
    if (hipMemsetAsync(p, 0, bytes, c->stream) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        (void)hipFree(p);
        return lr_fail(LRH_EDEVICE, "lrh_buffer_alloc: the buffer could not be cleared");
    }
    *d_out = p;
    return 0;
}

int lrh_buffer_free(LrhContext *c, void *d_buf)
{
    if (!c) return lr_fail(LRH_EINVAL, "lrh_buffer_free: no context");
    if (!d_buf) return 0;
    LR_HIP(hipSetDevice(c->device));
    LR_HIP(hipStreamSynchronize(c->stream));     // no work of the context still uses it
    LR_HIP(hipFree(d_buf));
    return 0;
}

int lrh_buffer_write(LrhContext *c, void *d_dst, const void *h_src, size_t bytes)
{
    if (!c || (bytes && (!d_dst || !h_src))) return lr_fail(LRH_EINVAL, "lrh_buffer_write: bad arguments");
    if (!bytes) return 0;
    LR_HIP(hipSetDevice(c->device));
    LR_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, c->stream));
    LR_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int lrh_buffer_read(LrhContext *c, void *h_dst, const void *d_src, size_t bytes)
{
    if (!c || (bytes && (!d_src || !h_dst))) return lr_fail(LRH_EINVAL, "lrh_buffer_read: bad arguments");
    if (!bytes) return 0;
    LR_HIP(hipSetDevice(c->device));
    // after the context's queued work (the kernels that write it), on its stream
    LR_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
    LR_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int lrh_tri_boxes(LrhContext *c, const float *d_tris, int n, float *d_boxes)
{
    if (!c || n < 0 || (n > 0 && (!d_tris || !d_boxes))) return lr_fail(LRH_EINVAL, "lrh_tri_boxes: bad arguments");
    return lr_launch_tri_boxes(c, d_tris, n, d_boxes);
}

int lrh_bounds(LrhContext *c, const float *d_boxes, int n, float *d_bounds)
{
    if (!c || n <= 0 || !d_boxes || !d_bounds) return lr_fail(LRH_EINVAL, "lrh_bounds: bad arguments");
    int rc = lr_scratch_reserve(c, LR_SCRATCH_BUILD, lr_align(sizeof(LrBuildParamsDev)) + lr_bounds_tmp_bytes());
    if (rc) return rc;
    char *S = (char *)c->scratch[LR_SCRATCH_BUILD].base;
    LrBuildParamsDev *prm = (LrBuildParamsDev *)S;
    rc = lr_launch_bounds(c, d_boxes, n, prm, S + lr_align(sizeof(LrBuildParamsDev)));
    if (rc) return rc;
    LR_HIP(hipMemcpyAsync(d_bounds, prm->lo, 12, hipMemcpyDeviceToDevice, c->stream));
    LR_HIP(hipMemcpyAsync(d_bounds + 3, prm->hi, 12, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

int lrh_build(LrhContext *c, const float *d_boxes, int n, const float *h_bounds, int builder, uint32_t node0,
              uint32_t ref0, uint32_t ref_add, uint32_t *d_nodes, uint32_t *d_aux, uint32_t *d_sorted,
              LrhBuildStats *stats)
{
    if (!c || !d_boxes || !d_nodes || !d_aux) return lr_fail(LRH_EINVAL, "lrh_build: NULL argument");
    if (n <= 0) return lr_fail(LRH_EINVAL, "lrh_build: %d items (nothing to build)", n);
    if (n > (1 << 29)) return lr_fail(LRH_EREFUSED, "lrh_build: %d items: at most 2^29", n);
    if (builder != LRH_LBVH && builder != LRH_PLOC) return lr_fail(LRH_EINVAL, "lrh_build: builder %d", builder);
    return lr_build_run(c, d_boxes, n, h_bounds, builder, node0, ref0, ref_add, d_nodes, d_aux, d_sorted,
                        d_sorted ? d_sorted + n : nullptr, stats);
}

int lrh_build_noalloc(LrhContext *c, const float *d_boxes, int n, const float *h_bounds, int builder, uint32_t node0,
                      uint32_t ref0, uint32_t ref_add, uint32_t *d_nodes, uint32_t *d_aux, uint32_t *d_sorted,
                      LrhBuildStats *stats)
{
    if (!c || !d_boxes || !d_nodes || !d_aux) return lr_fail(LRH_EINVAL, "lrh_build_noalloc: NULL argument");
    if (n <= 0) return lr_fail(LRH_EINVAL, "lrh_build_noalloc: %d items (nothing to build)", n);
    if (n > LRH_NOALLOC_MAX_ITEMS)
        return lr_fail(LRH_EREFUSED, "lrh_build_noalloc: %d items: at most %d (its scratch is fixed)", n,
                       LRH_NOALLOC_MAX_ITEMS);
    if (builder != LRH_LBVH && builder != LRH_PLOC) return lr_fail(LRH_EINVAL, "lrh_build_noalloc: builder %d", builder);
    LR_HIP(hipSetDevice(c->device));
    return lr_build_noalloc(c, d_boxes, n, h_bounds, builder, node0, ref0, ref_add, d_nodes, d_aux, d_sorted, stats);
}

int lrh_trace(LrhContext *c, const uint32_t *d_nodes, uint32_t root, const uint32_t *d_aux, const float *d_tris,
              const float *d_rays, int n, int any, float *d_hits, int max_depth)
{
    if (!c || n < 0 || (n > 0 && (!d_nodes || !d_aux || !d_tris || !d_rays || !d_hits)))
        return lr_fail(LRH_EINVAL, "lrh_trace: bad arguments");
    return lr_launch_trace(c, d_nodes, root, d_aux, d_tris, d_rays, n, nullptr, any, d_hits, max_depth);
}

int lrh_trace_queue(LrhContext *c, const uint32_t *d_nodes, uint32_t root, const uint32_t *d_aux, const float *d_tris,
                    const float *d_rays, int max_n, const uint32_t *d_count, int any, float *d_hits, int max_depth)
{
    if (!c || max_n < 0 || !d_count || (max_n > 0 && (!d_nodes || !d_aux || !d_tris || !d_rays || !d_hits)))
        return lr_fail(LRH_EINVAL, "lrh_trace_queue: bad arguments");
    return lr_launch_trace(c, d_nodes, root, d_aux, d_tris, d_rays, max_n, d_count, any, d_hits, max_depth);
}

int lrh_trace_count(LrhContext *c, const uint32_t *d_nodes, uint32_t root, const uint32_t *d_aux, const float *d_tris,
                    const float *d_rays, int max_n, const uint32_t *d_count, int any, float *d_hits, int max_depth,
                    uint64_t *d_counts)
{
    if (!c || max_n < 0 || !d_counts || (max_n > 0 && (!d_nodes || !d_aux || !d_tris || !d_rays || !d_hits)))
        return lr_fail(LRH_EINVAL, "lrh_trace_count: bad arguments");
    return lr_launch_trace(c, d_nodes, root, d_aux, d_tris, d_rays, max_n, d_count, any, d_hits, max_depth,
                           (unsigned long long *)d_counts);
}

int lrh_rays_primary(LrhContext *c, const float *cam, int w, int h, float tmin, float tmax, float *d_rays)
{
    if (!c || !cam || w <= 0 || h <= 0 || !d_rays) return lr_fail(LRH_EINVAL, "lrh_rays_primary: bad arguments");
    return lr_launch_rays_primary(c, cam, w, h, tmin, tmax, d_rays);
}

int lrh_rays_secondary(LrhContext *c, const float *d_rays, const float *d_hits, int n, const float *d_tris,
                       const float *light, float eps, uint32_t seed, float *d_bounce, float *d_shadow, uint32_t *d_pix,
                       uint32_t *d_count)
{
    if (!c || n < 0 || !light || !d_count || (n > 0 && (!d_rays || !d_hits || !d_tris || !d_bounce || !d_shadow || !d_pix)))
        return lr_fail(LRH_EINVAL, "lrh_rays_secondary: bad arguments");
    return lr_launch_rays_secondary(c, d_rays, d_hits, n, d_tris, light, eps, seed, d_bounce, d_shadow, d_pix, d_count);
}

}  // extern "C"
