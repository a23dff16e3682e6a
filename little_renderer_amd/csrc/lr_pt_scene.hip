// lr_pt_scene.hip - scenes resident on the device for the path tracer (lrh_pt_scene): the
// shapes, materials and BVH of one Scene (engine/scene/scene.h:300-331), uploaded once and
// kept until freed, so that any number of jobs can render them.
#include <string.h>

#include <vector>

#include "lr_device.h"
#include "lr_engine.h"

// The engine's node records (walk2, lr_walk.h): for each inner node of the reference layout,
// both children's boxes and (index | leaf record). A one-shape tree (its root a leaf) gets a
// record whose left child is that leaf and whose right child is an empty leaf.
__global__ __launch_bounds__(256) void k_child_boxes(const float4 *__restrict__ nodes, int nn, float4 *__restrict__ w)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    const float4 lo = nodes[2 * (size_t)i], hi = nodes[2 * (size_t)i + 1];
    const uint32_t b = __float_as_uint(hi.w);
    if (b & LR_LEAF) {
        if (i != 0) return;
        w[0] = lo;
        w[1] = hi;
        w[2] = lo;
        w[3] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(LR_LEAF));
        return;
    }
    const uint32_t c[2] = {__float_as_uint(lo.w), b};
    for (int k = 0; k < 2; k++) {
        float4 clo = nodes[2 * (size_t)c[k]], chi = nodes[2 * (size_t)c[k] + 1];
        if (!(__float_as_uint(chi.w) & LR_LEAF)) {
            clo.w = __uint_as_float(c[k]);
            chi.w = __uint_as_float(0u);
        }
        w[4 * (size_t)i + 2 * k] = clo;
        w[4 * (size_t)i + 2 * k + 1] = chi;
    }
}

static void pt_scene_release(LrhPtScene *s)
{
    if (!s) return;
    void *ps[] = {s->prims, s->tris, s->mats, s->boxes, s->nodes, s->aux, s->tab, s->mask, s->wnodes};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    delete s;
}

extern "C" void lrs_blue_noise(uint16_t *rank);
extern "C" void lrs_brdf_tables(float *out);

static void quad_bounds(const float *p, float *lo, float *hi)
{
    // primitive_bounds (scene.c:47-63): the 4 corners (or the sphere's box), padded by 1e-3
    if (p[4 + 3] < 0.5f) {
        for (int k = 0; k < 3; k++) { lo[k] = p[k] - p[3]; hi[k] = p[k] + p[3]; }
    } else {
        const float *q = p + 4, *u = p + 8, *v = p + 12;
        for (int k = 0; k < 3; k++) {
            const float c[4] = {q[k], q[k] + u[k], q[k] + v[k], q[k] + (u[k] + v[k])};
            lo[k] = hi[k] = c[0];
            for (int j = 1; j < 4; j++) {
                lo[k] = c[j] != c[j] ? lo[k] : (lo[k] < c[j] ? lo[k] : c[j]);
                hi[k] = c[j] != c[j] ? hi[k] : (hi[k] > c[j] ? hi[k] : c[j]);
            }
        }
    }
    for (int k = 0; k < 3; k++) { lo[k] -= 1e-3f; hi[k] += 1e-3f; }
}

extern "C" int lrh_pt_scene(LrhContext *c, const float *prims, int nprims, const float *tris, int ntris,
                            const LrhMaterial *mats, int nmats, int light, float eps, float margin, int builder,
                            LrhPtScene **out)
{
    if (!c || !out || nprims < 0 || ntris < 0 || nprims + ntris <= 0 || nmats <= 0 || nmats > 64 || !mats ||
        (nprims > 0 && !prims) || (ntris > 0 && !tris) || light >= nprims)
        return lr_fail(LRH_EINVAL, "lrh_pt_scene: bad arguments");
    if (light >= 0 && prims[24 * light + 7] < 0.5f) return lr_fail(LRH_EINVAL, "lrh_pt_scene: the lamp must be a quad");
    for (int i = 0; i < nprims; i++) {
        const float m = prims[24 * (size_t)i + 11];
        if (!(m >= 0.0f && m < (float)nmats)) return lr_fail(LRH_EINVAL, "lrh_pt_scene: prim %d: material %g of %d", i, m, nmats);
    }
    for (int i = 0; i < ntris; i++) {
        uint32_t w;
        memcpy(&w, tris + 12 * (size_t)i + 11, 4);
        if ((int)(w >> 26) >= nmats) return lr_fail(LRH_EINVAL, "lrh_pt_scene: triangle %d: material %u of %d", i, w >> 26, nmats);
    }
    for (int i = 0; i < nmats; i++)
        if (mats[i].type < 0 || mats[i].type > LRH_MAT_PBR) return lr_fail(LRH_EINVAL, "lrh_pt_scene: material %d: type %d", i, mats[i].type);
    *out = nullptr;
    LrhPtScene *s = new LrhPtScene();
    s->ctx = c;
    s->nprims = nprims;
    s->ntris = ntris;
    s->nmats = nmats;
    s->light = light;
    s->eps = eps;
    s->margin = margin;
    const int n = nprims + ntris;
    std::vector<float> hm(12 * (size_t)nmats), boxes(8 * (size_t)n, 0.0f);
    for (int i = 0; i < nmats; i++) {
        const LrhMaterial &m = mats[i];
        float *o = &hm[12 * (size_t)i];
        o[0] = m.albedo[0]; o[1] = m.albedo[1]; o[2] = m.albedo[2]; o[3] = (float)m.type;
        o[4] = m.emission[0]; o[5] = m.emission[1]; o[6] = m.emission[2]; o[7] = m.ior;
        o[8] = m.metallic; o[9] = m.roughness; o[10] = m.transmission; o[11] = 0.0f;
    }
    for (int i = 0; i < nprims; i++) quad_bounds(prims + 24 * (size_t)i, &boxes[8 * (size_t)i], &boxes[8 * (size_t)i + 4]);
    for (int i = 0; i < ntris; i++) {
        const float *t = tris + 12 * (size_t)i;
        float *o = &boxes[8 * (size_t)(nprims + i)];
        for (int k = 0; k < 3; k++) {   // triangle_bounds (scene.c:65-74)
            const float q = t[k], a = q + t[4 + k], b = q + t[8 + k];
            float lo = q, hi = q;
            lo = a != a ? lo : (lo < a ? lo : a); hi = a != a ? hi : (hi > a ? hi : a);
            lo = b != b ? lo : (lo < b ? lo : b); hi = b != b ? hi : (hi > b ? hi : b);
            o[k] = lo - 1e-3f;
            o[4 + k] = hi + 1e-3f;
        }
    }
    std::vector<float> tab(2112);
    std::vector<uint16_t> mask(4096);
    lrs_brdf_tables(tab.data());
    lrs_blue_noise(mask.data());
    auto up = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0) bytes = 16;
        if (hipMalloc(dst, bytes) != hipSuccess) return false;
        if (src && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return false;
        return true;
    };
    bool ok = up((void **)&s->prims, nprims ? prims : nullptr, (size_t)nprims * 96) &&
              up((void **)&s->tris, ntris ? tris : nullptr, (size_t)ntris * 48) &&
              up((void **)&s->mats, hm.data(), hm.size() * 4) && up((void **)&s->boxes, boxes.data(), boxes.size() * 4) &&
              up((void **)&s->nodes, nullptr, (2 * (size_t)n - 1) * 32) && up((void **)&s->aux, nullptr, (size_t)n * 4) &&
              up((void **)&s->tab, tab.data(), tab.size() * 4) && up((void **)&s->mask, mask.data(), mask.size() * 2);
    if (!ok) {
        (void)hipGetLastError();
        pt_scene_release(s);
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene: device memory for the scene could not be had");
    }
    LrhBuildStats st;
    int rc = lrh_build(c, (const float *)s->boxes, n, nullptr, builder, 0, 0, 0, s->nodes, s->aux, nullptr, &st);
    if (rc) {
        pt_scene_release(s);
        return rc;
    }
    s->depth = st.depth;
    const size_t nn = 2 * (size_t)n - 1;
    if (hipMalloc((void **)&s->wnodes, nn * 64) != hipSuccess) {
        (void)hipGetLastError();
        pt_scene_release(s);
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene: device memory for the scene could not be had");
    }
    hipLaunchKernelGGL(k_child_boxes, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream,
                       (const float4 *)s->nodes, (int)nn, s->wnodes);
    s->bytes = (size_t)nprims * 96 + (size_t)ntris * 48 + hm.size() * 4 + boxes.size() * 4 + (2 * (size_t)n - 1) * 32 +
               (size_t)n * 4 + tab.size() * 4 + mask.size() * 2 + (2 * (size_t)n - 1) * 64;
    // the tree is built on the context's stream; the engines that render the scene run on
    // streams of their own, so the scene is complete before it is handed out
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        (void)hipGetLastError();
        pt_scene_release(s);
        return lr_fail(LRH_EDEVICE, "lrh_pt_scene: the build failed on the device");
    }
    *out = s;
    return 0;
}

extern "C" void lrh_pt_scene_free(LrhPtScene *s)
{
    if (s && s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    pt_scene_release(s);
}

extern "C" int lrh_pt_scene_info(const LrhPtScene *s, int *depth, int *nshapes, size_t *bytes)
{
    if (!s) return lr_fail(LRH_EINVAL, "lrh_pt_scene_info: no scene");
    if (depth) *depth = s->depth;
    if (nshapes) *nshapes = s->nprims + s->ntris;
    if (bytes) *bytes = s->bytes;
    return 0;
}
