// lr_pt_scene.hip - scenes resident on the device for the path tracer (lrh_pt_scene): the
// shapes, materials and BVH of one Scene (engine/scene/scene.h:300-331), uploaded once and
// kept until freed, so that any number of jobs can render them.
#include <string.h>

#include <vector>

#include "lr_device.h"
#include "lr_engine_k.h"

// Binary records with both children's boxes in the parent: for each inner node of the reference
// layout, both children's boxes and (index | leaf record); a one-shape tree (its root a leaf)
// gets a record whose left child is that leaf and whose right child is an empty leaf. Made for
// the two-level scenes' top level and meshes (JobDev bnodes, mnodes); no walk reads them now.
__global__ __launch_bounds__(256) void k_child_boxes(const float4 *__restrict__ nodes, int i0, int nn, float4 *__restrict__ w)
{
    const int i = i0 + (int)(blockIdx.x * 256 + threadIdx.x);   // records i0 .. i0 + nn - 1
    if (i >= i0 + nn) return;
    const float4 lo = nodes[2 * (size_t)i], hi = nodes[2 * (size_t)i + 1];
    const uint32_t b = __float_as_uint(hi.w);
    if (b & LR_LEAF) return;
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

// The triangles in leaf order: entry e of aux names shape aux[e]; a triangle's 48 bytes go to
// ltris[3 e..], a primitive's entry gets zeros (the walks test primitives from prims).
__global__ __launch_bounds__(256) void k_leaf_tris(const uint32_t *__restrict__ aux, int n, uint32_t nprims,
                                                   const float4 *__restrict__ tris, float4 *__restrict__ ltris)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const uint32_t sid = aux[e];
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a, c = a;
    if (sid >= nprims) {
        const size_t t = sid - nprims;
        a = tris[3 * t];
        b = tris[3 * t + 1];
        c = tris[3 * t + 2];
    }
    ltris[3 * (size_t)e] = a;
    ltris[3 * (size_t)e + 1] = b;
    ltris[3 * (size_t)e + 2] = c;
}

// A tree whose root is a leaf (one shape) gets a record with that leaf on the left and an
// empty leaf on the right.
__global__ void k_root_leaf(const float4 *__restrict__ nodes, uint32_t r, float4 *__restrict__ w)
{
    const float4 lo = nodes[2 * (size_t)r], hi = nodes[2 * (size_t)r + 1];
    if (!(__float_as_uint(hi.w) & LR_LEAF)) return;
    w[4 * (size_t)r] = lo;
    w[4 * (size_t)r + 1] = hi;
    w[4 * (size_t)r + 2] = lo;
    w[4 * (size_t)r + 3] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(LR_LEAF));
}

// Each instance's world box: the 8 corners of its object box (its mesh's root box, or [-1, 1]^3)
// moved to the world, padded by a little more than their rounding.
__global__ __launch_bounds__(256) void k_inst_boxes(const float4 *__restrict__ fwd, const float4 *__restrict__ inst,
                                                    const uint32_t *__restrict__ mroot, const float4 *__restrict__ mnodes,
                                                    int n, float4 *__restrict__ boxes)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t kind = __float_as_uint(inst[4 * (size_t)i + 3].x);
    float lo[3] = {-1.0f, -1.0f, -1.0f}, hi[3] = {1.0f, 1.0f, 1.0f};
    if (kind < LR_UNIT_SPHERE) {
        const uint32_t r = mroot[kind];
        const float4 a = mnodes[2 * (size_t)r], b = mnodes[2 * (size_t)r + 1];
        lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
        hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    }
    const float4 rows[3] = {fwd[3 * (size_t)i], fwd[3 * (size_t)i + 1], fwd[3 * (size_t)i + 2]};
    float wlo[3] = {1e30f, 1e30f, 1e30f}, whi[3] = {-1e30f, -1e30f, -1e30f};
    for (int c = 0; c < 8; c++) {
        const float p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
        for (int k = 0; k < 3; k++) {
            const float4 r = rows[k];
            const float v = r.x * p[0] + r.y * p[1] + r.z * p[2] + r.w;
            wlo[k] = fminf(wlo[k], v);
            whi[k] = fmaxf(whi[k], v);
        }
    }
    float4 bl, bh;
    float *ol = &bl.x, *oh = &bh.x;
    for (int k = 0; k < 3; k++) {
        const float pad = 1e-5f * fmaxf(whi[k] - wlo[k], fmaxf(fabsf(wlo[k]), fabsf(whi[k]))) + 1e-6f;
        ol[k] = wlo[k] - pad;
        oh[k] = whi[k] + pad;
    }
    bl.w = 0.0f;
    bh.w = 0.0f;
    boxes[2 * (size_t)i] = bl;
    boxes[2 * (size_t)i + 1] = bh;
}

// ---- the 4-wide tree (bvh_make_wide / collapse, engine/rt/bvh.c:487-539): level by level, one
// wide node per thread. The children are chosen as the reference chooses them; the node is
// stored compressed (64 bytes instead of put_wide's four 32-byte slots, trace_gpu.c:867-881):
//   float4 0: the node's box corner lo.xyz, then the three axes' scale exponents (8 bits each,
//             biased by 127: the scale is 2^(e - 127))
//   float4 1: the four children's references: a wide node's index; a leaf 0x80000000 |
//             (count - 1) << 24 | its first aux entry; no child LR_NONE
//   float4 2: the children's lo.x, lo.y, lo.z, hi.x as 8-bit steps from the corner (byte k:
//             child k), float4 3: hi.y, hi.z, 0, 0
// Each child's box is rounded outwards to the steps (lo down, hi up, in exact arithmetic), so it
// holds the child's own box: the walk visits a few more nodes, never misses a hit. ----
struct WideTask {
    uint32_t node, wide;   // the binary node the wide node stands for, its index
};

__device__ __forceinline__ float node_area(const float4 *nodes, uint32_t i)   // box_area (bvh.c:30-34), halved
{
    const float4 lo = nodes[2 * (size_t)i], hi = nodes[2 * (size_t)i + 1];
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
}

__global__ void k_collapse_init(WideTask *q0, uint32_t *cnt, int levels, uint32_t *nwide, uint32_t root, uint32_t w0)
{
    for (int l = 0; l <= levels + 1; l++) cnt[l] = 0;
    q0[0] = WideTask{root, w0};
    cnt[0] = 1;
    *nwide = w0 + 1;
}

__global__ __launch_bounds__(256) void k_collapse(const float4 *__restrict__ nodes, const WideTask *__restrict__ qin,
                                                  const uint32_t *cin, WideTask *__restrict__ qout, uint32_t *cout,
                                                  uint32_t *nwide, float4 *__restrict__ w4)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= *cin) return;
    const WideTask e = qin[i];
    uint32_t kids[4];
    int n = 0;
    const uint32_t rb = __float_as_uint(nodes[2 * (size_t)e.node + 1].w);
    if (rb & LR_LEAF) kids[n++] = e.node;   // a tree that is one leaf
    else {
        kids[n++] = __float_as_uint(nodes[2 * (size_t)e.node].w);
        kids[n++] = rb;
    }
    // open the inner child of largest area until there are four (or only leaves)
    while (n < 4) {
        int best = -1;
        float area = -1.0f;
        for (int k = 0; k < n; k++) {
            const bool inner = !(__float_as_uint(nodes[2 * (size_t)kids[k] + 1].w) & LR_LEAF);
            const float a = node_area(nodes, kids[k]);
            if (inner && a > area) {
                area = a;
                best = k;
            }
        }
        if (best < 0) break;
        const uint32_t c = kids[best];
        kids[best] = __float_as_uint(nodes[2 * (size_t)c].w);
        kids[n++] = __float_as_uint(nodes[2 * (size_t)c + 1].w);
    }
    // the node's box (the union of its children's) and, per axis, the smallest power of two s
    // with extent / s < 255
    const float4 plo = nodes[2 * (size_t)e.node], phi = nodes[2 * (size_t)e.node + 1];
    const float pl[3] = {plo.x, plo.y, plo.z}, ph[3] = {phi.x, phi.y, phi.z};
    int ex[3];
    for (int a = 0; a < 3; a++) {
        const double ext = (double)ph[a] - (double)pl[a];
        int k = -126;
        if (ext > 0.0) {
            (void)frexp(ext / 255.0, &k);   // ext / 255 < 2^k
            k = k < -126 ? -126 : (k > 127 ? 127 : k);
        }
        ex[a] = k;
    }
    uint32_t ref[4] = {LR_NONE, LR_NONE, LR_NONE, LR_NONE}, qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
    for (int k = 0; k < n; k++) {
        const float4 lo = nodes[2 * (size_t)kids[k]], hi = nodes[2 * (size_t)kids[k] + 1];
        const uint32_t b = __float_as_uint(hi.w);
        if (b & LR_LEAF) ref[k] = LR_LEAF | (((b & 0x7fffffffu) - 1u) << 24) | __float_as_uint(lo.w);
        else {
            const uint32_t w = atomicAdd(nwide, 1u);
            qout[atomicAdd(cout, 1u)] = WideTask{kids[k], w};
            ref[k] = w;
        }
        const float cl[3] = {lo.x, lo.y, lo.z}, ch[3] = {hi.x, hi.y, hi.z};
        for (int a = 0; a < 3; a++) {
            // steps from the corner, exact in double (a float difference, a power-of-two scale)
            const double fl = floor(ldexp((double)cl[a] - (double)pl[a], -ex[a]));
            const double fh = ceil(ldexp((double)ch[a] - (double)pl[a], -ex[a]));
            const uint32_t l = fl < 0.0 ? 0u : (fl > 255.0 ? 255u : (uint32_t)fl);
            const uint32_t h = fh < 0.0 ? 0u : (fh > 255.0 ? 255u : (uint32_t)fh);
            qlo[a] |= l << (8 * k);
            qhi[a] |= h << (8 * k);
        }
    }
    const uint32_t eb = (uint32_t)(ex[0] + 127) | ((uint32_t)(ex[1] + 127) << 8) | ((uint32_t)(ex[2] + 127) << 16);
    float4 *o = w4 + 4 * (size_t)e.wide;
    o[0] = make_float4(plo.x, plo.y, plo.z, __uint_as_float(eb));
    o[1] = make_float4(__uint_as_float(ref[0]), __uint_as_float(ref[1]), __uint_as_float(ref[2]), __uint_as_float(ref[3]));
    o[2] = make_float4(__uint_as_float(qlo[0]), __uint_as_float(qlo[1]), __uint_as_float(qlo[2]), __uint_as_float(qhi[0]));
    o[3] = make_float4(__uint_as_float(qhi[1]), __uint_as_float(qhi[2]), 0.0f, 0.0f);
}

// the wide tree of a binary tree (root `root`, n items, depth levels) into w4, its wide nodes
// numbered from w0 (room for n of them from there); waits for it
static int collapse_wide(LrhContext *c, const uint32_t *nodes, uint32_t root, int n, int depth, float4 *w4, uint32_t w0 = 0)
{
    const size_t qbytes = lr_align((size_t)(n > 1 ? n : 2) * sizeof(WideTask));
    int rc = lr_scratch_reserve(c, LR_SCRATCH_TRACE, 2 * qbytes + lr_align((size_t)(depth + 4) * 4) + 256);
    if (rc) return rc;
    char *base = (char *)c->scratch[LR_SCRATCH_TRACE].base;
    WideTask *q[2] = {(WideTask *)base, (WideTask *)(base + qbytes)};
    uint32_t *cnt = (uint32_t *)(base + 2 * qbytes);
    uint32_t *nwide = cnt + depth + 3;
    hipLaunchKernelGGL(k_collapse_init, dim3(1), dim3(1), 0, c->stream, q[0], cnt, depth, nwide, root, w0);
    const dim3 grid((unsigned)((n + 255) / 256));
    for (int l = 0; l <= depth; l++)
        hipLaunchKernelGGL(k_collapse, grid, dim3(256), 0, c->stream, (const float4 *)nodes, (const WideTask *)q[l & 1],
                           (const uint32_t *)(cnt + l), q[(l + 1) & 1], cnt + l + 1, nwide, w4);
    LR_HIP(hipGetLastError());
    LR_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

static void pt_scene_release(LrhPtScene *s)
{
    if (!s) return;
    void *ps[] = {s->prims, s->tris, s->mats, s->boxes, s->nodes, s->aux, s->tab, s->mask, s->wnodes, s->w4nodes, s->ltris, s->env_rgb,
                  s->env_marg, s->env_cond, s->env_prob, s->mboxes, s->mwnodes, s->inst, s->ifwd, s->mnodes, s->maux,
                  s->mroot, s->mw4nodes, s->tw4nodes, s->mroot4, s->tparent, s->tarrive, s->tcost_dev, s->emit, s->inst_emit};
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
    if ((int64_t)nprims + ntris > LR_WIDE_MAX_SHAPES)
        return lr_fail(LRH_EREFUSED, "lrh_pt_scene: %d shapes, the wide nodes' leaf references take at most %d",
                       nprims + ntris, LR_WIDE_MAX_SHAPES);
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
    s->hmats.assign(mats, mats + nmats);
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
    // the 4-wide tree the jobs walk; its stack: up to 3 entries per wide level
    if (hipMalloc((void **)&s->w4nodes, (size_t)n * 64) != hipSuccess) {
        (void)hipGetLastError();
        pt_scene_release(s);
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene: device memory for the scene could not be had");
    }
    rc = collapse_wide(c, s->nodes, 0u, n, st.depth, s->w4nodes);
    if (rc) {
        pt_scene_release(s);
        return rc;
    }
    // the triangles in leaf order: a leaf's triangle is then one load after its node, beside
    // (not after) its aux entry
    if (hipMalloc((void **)&s->ltris, (size_t)n * 48) != hipSuccess) {
        (void)hipGetLastError();
        pt_scene_release(s);
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene: device memory for the scene could not be had");
    }
    hipLaunchKernelGGL(k_leaf_tris, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, s->aux, n, (uint32_t)nprims,
                       (const float4 *)s->tris, s->ltris);
    s->depth = 2 * st.depth + 4;
    s->bytes = (size_t)nprims * 96 + (size_t)ntris * 48 + hm.size() * 4 + boxes.size() * 4 + (2 * (size_t)n - 1) * 32 +
               (size_t)n * 4 + tab.size() * 4 + mask.size() * 2 + (size_t)n * 64 + (size_t)n * 48;
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

extern "C" int lrs_sky_tables(const float *rgb, int w, int h, float *marg, float *cond, float *prob, double *sum);
extern "C" int lrs_affine_inverse(const float fwd[12], float inv[12]);

static bool dev_upload(void **dst, const void *src, size_t bytes)
{
    if (hipMalloc(dst, bytes ? bytes : 16) != hipSuccess) return false;
    if (src && bytes && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return false;
    return true;
}

extern "C" int lrh_pt_scene_sky(LrhPtScene *s, const float *rgb, int w, int h, float scale, float rot)
{
    if (!s || (rgb && (w < 2 || h < 1 || w != 2 * h)) || !(scale >= 0.0f))
        return lr_fail(LRH_EINVAL, "lrh_pt_scene_sky: bad arguments (the map is w x w/2)");
    (void)hipStreamSynchronize(s->ctx->stream);
    void *old[] = {s->env_rgb, s->env_marg, s->env_cond, s->env_prob};
    for (void *p : old)
        if (p) (void)hipFree(p);
    s->env_rgb = s->env_marg = s->env_cond = s->env_prob = nullptr;
    s->env_w = s->env_h = 0;
    if (!rgb) return 0;
    std::vector<float> marg(h + 1), cond((size_t)h * (w + 1)), prob((size_t)w * h);
    double sum = 0;
    if (lrs_sky_tables(rgb, w, h, marg.data(), cond.data(), prob.data(), &sum) != 0)
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene_sky: host memory for the tables could not be had");
    if (!dev_upload((void **)&s->env_rgb, rgb, (size_t)w * h * 12) || !dev_upload((void **)&s->env_marg, marg.data(), marg.size() * 4) ||
        !dev_upload((void **)&s->env_cond, cond.data(), cond.size() * 4) ||
        !dev_upload((void **)&s->env_prob, prob.data(), prob.size() * 4)) {
        (void)hipGetLastError();
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene_sky: device memory for the sky could not be had");
    }
    s->env_w = w;
    s->env_h = h;
    s->env_scale = scale;
    s->env_rot = rot;
    s->bytes += (size_t)w * h * 12 + marg.size() * 4 + cond.size() * 4 + prob.size() * 4;
    return 0;
}

// the instances' device records (world -> object rows, then kind, material, mesh root) and
// their object -> world rows, from the caller's transforms
static int pack_instances(const LrhPtScene *s, const LrhInstance *in, int n, const std::vector<uint32_t> &roots,
                          const std::vector<uint32_t> &roots4, std::vector<float> &rec, std::vector<float> &fwd)
{
    rec.assign((size_t)n * 16, 0.0f);
    fwd.assign((size_t)n * 12, 0.0f);
    for (int i = 0; i < n; i++) {
        const LrhInstance &x = in[i];
        if (x.kind < LRH_INST_BOX || x.kind >= s->nmeshes) return lr_fail(LRH_EINVAL, "instance %d: kind %d", i, x.kind);
        if (x.material < -1 || x.material >= s->nmats) return lr_fail(LRH_EINVAL, "instance %d: material %d", i, x.material);
        float inv[12];
        if (lrs_affine_inverse(x.xform, inv) != 0) return lr_fail(LRH_EINVAL, "instance %d: its transform is singular", i);
        float *o = &rec[(size_t)i * 16];
        memcpy(o, inv, 48);
        const uint32_t kind = x.kind == LRH_INST_SPHERE ? LR_UNIT_SPHERE : x.kind == LRH_INST_BOX ? LR_UNIT_BOX : (uint32_t)x.kind;
        const uint32_t mat = x.material < 0 ? LR_NONE : (uint32_t)x.material;
        const uint32_t root = x.kind >= 0 ? roots[x.kind] : 0u, root4 = x.kind >= 0 ? roots4[x.kind] : 0u;
        memcpy(o + 12, &kind, 4);
        memcpy(o + 13, &mat, 4);
        memcpy(o + 14, &root, 4);
        memcpy(o + 15, &root4, 4);
        memcpy(&fwd[(size_t)i * 12], x.xform, 48);
    }
    return 0;
}

// ---- the glowing shapes (Scene.emitters): scene_find_emitters and scene_emitter
// (engine/scene/scene.c:1000-1031) over the host copies of the instances and materials ----
static bool material_glows(const LrhMaterial &m)   // scene.h:87-89
{
    return m.type != LRH_MAT_LIGHT && (m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f);
}
static float len3(const float *a) { return sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static void cross3(const float *a, const float *b, float *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// the emitter table (5 float4 each: centre, ball | half axis 0, radius | half axis 1, area |
// half axis 2 | radiance) and each instance's slot, made again from the current poses
static int emitters_again(LrhPtScene *s)
{
    std::vector<float> rec(LR_MAX_EMITTERS * 20, 0.0f);
    std::vector<int> slot((size_t)s->ninst, -1);
    int n = 0;
    for (int i = 0; i < s->ninst && n < LR_MAX_EMITTERS; i++) {
        const LrhInstance &x = s->hinst[(size_t)i];
        if ((x.kind != LRH_INST_SPHERE && x.kind != LRH_INST_BOX) || x.material < 0 || x.material >= s->nmats) continue;
        const LrhMaterial &m = s->hmats[(size_t)x.material];
        if (!material_glows(m)) continue;
        float a[3][3];   // the transform's columns: the shape's half axes in the world
        for (int k = 0; k < 3; k++)
            for (int r = 0; r < 3; r++) a[k][r] = x.xform[4 * r + k];
        const bool ball = x.kind == LRH_INST_SPHERE;
        if (ball) {      // a stretched ball is no light
            const float la = len3(a[0]), lb = len3(a[1]), lc = len3(a[2]);
            if (!(la > 0.0f) || fabsf(la - lb) > 1e-4f * la || fabsf(la - lc) > 1e-4f * la) continue;
        }
        float *o = &rec[(size_t)n * 20];
        o[0] = x.xform[3];
        o[1] = x.xform[7];
        o[2] = x.xform[11];
        o[3] = ball ? 1.0f : 0.0f;
        for (int k = 0; k < 3; k++)
            for (int r = 0; r < 3; r++) o[4 + 4 * k + r] = a[k][r];
        if (ball) {
            const float r = len3(a[0]);
            o[7] = r;
            o[11] = 4.0f * PI_F * r * r;
        } else {
            float c12[3], c02[3], c01[3];
            cross3(a[1], a[2], c12);
            cross3(a[0], a[2], c02);
            cross3(a[0], a[1], c01);
            o[11] = 8.0f * (len3(c12) + len3(c02) + len3(c01));
        }
        for (int k = 0; k < 3; k++) o[16 + k] = m.emission[k];
        slot[(size_t)i] = n++;
    }
    LR_HIP(hipStreamSynchronize(s->ctx->stream));
    LR_HIP(hipMemcpy(s->emit, rec.data(), rec.size() * 4, hipMemcpyHostToDevice));
    LR_HIP(hipMemcpy(s->inst_emit, slot.data(), slot.size() * 4, hipMemcpyHostToDevice));
    s->nemit = n;
    return 0;
}

// ---- the top level refitted (LRH_REFIT: bvh_refit_boxes and the refit rule of build_tlas,
// engine/rt/bvh.c:464-483, engine/rt/accel.c:152-173): the same tree over the instances' new
// boxes while its cost (the nodes' areas over the root's) stays within 1.1 of its last build's ----
constexpr double LR_REFIT_GROWTH = 1.1;

// each record's parent (the root's: LR_NONE)
__global__ __launch_bounds__(256) void k_parents(const float4 *__restrict__ nodes, int nn, uint32_t *__restrict__ parent)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    if (i == 0) parent[0] = LR_NONE;
    const uint32_t a = __float_as_uint(nodes[2 * (size_t)i].w), b = __float_as_uint(nodes[2 * (size_t)i + 1].w);
    if (b & LR_LEAF) return;
    parent[a] = (uint32_t)i;
    parent[b] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_clear32(uint32_t *p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// From each leaf record (its box: its instances' world boxes) up: at each inner record the first
// of its two children's threads to arrive stops, the second joins their boxes (left, then right)
// into it and goes on. Records are handed over through the coherence point (sc1 stores, waited
// for before the arrival; sc1 loads after it), as the LBVH's climbs are.
__global__ __launch_bounds__(256) void k_refit(float4 *nodes, int nn, const uint32_t *__restrict__ aux,
                                               const float4 *__restrict__ boxes, const uint32_t *__restrict__ parent,
                                               uint32_t *arrive)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    const float4 r0 = nodes[2 * (size_t)i], r1 = nodes[2 * (size_t)i + 1];
    const uint32_t b = __float_as_uint(r1.w);
    if (!(b & LR_LEAF)) return;
    const uint32_t first = __float_as_uint(r0.w), cnt = b & 0x7fffffffu;
    float lx = 3e38f, ly = 3e38f, lz = 3e38f, hx = -3e38f, hy = -3e38f, hz = -3e38f;
    for (uint32_t k = 0; k < cnt; k++) {
        const uint32_t it = aux[first + k];
        const float4 blo = boxes[2 * (size_t)it], bhi = boxes[2 * (size_t)it + 1];
        lx = k ? wmin(lx, blo.x) : blo.x; ly = k ? wmin(ly, blo.y) : blo.y; lz = k ? wmin(lz, blo.z) : blo.z;
        hx = k ? wmax(hx, bhi.x) : bhi.x; hy = k ? wmax(hy, bhi.y) : bhi.y; hz = k ? wmax(hz, bhi.z) : bhi.z;
    }
    uint64_t *w = reinterpret_cast<uint64_t *>(nodes);
    auto put = [&](uint32_t x, uint32_t a, uint32_t bb) {
        st_dev64(w + 4 * (size_t)x, ((uint64_t)__float_as_uint(ly) << 32) | __float_as_uint(lx));
        st_dev64(w + 4 * (size_t)x + 1, ((uint64_t)a << 32) | __float_as_uint(lz));
        st_dev64(w + 4 * (size_t)x + 2, ((uint64_t)__float_as_uint(hy) << 32) | __float_as_uint(hx));
        st_dev64(w + 4 * (size_t)x + 3, ((uint64_t)bb << 32) | __float_as_uint(hz));
    };
    put((uint32_t)i, first, b);
    uint32_t cur = (uint32_t)i;
    for (;;) {
        const uint32_t p = parent[cur];
        if (p == LR_NONE) return;                   // the root
        // release: this record is visible before the arrival; acquire: the sibling's record,
        // written before its arrival, is visible to the loads below (at most 2 ninst - 1
        // records, so the guaranteed ordering costs nothing that matters)
        if (__hip_atomic_fetch_add(&arrive[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
        const uint32_t pa = __float_as_uint(nodes[2 * (size_t)p].w), pb = __float_as_uint(nodes[2 * (size_t)p + 1].w);
        const uint32_t sib = pa == cur ? pb : pa;
        const uint64_t q0 = ld_dev64(w + 4 * (size_t)sib), q1 = ld_dev64(w + 4 * (size_t)sib + 1);
        const uint64_t q2 = ld_dev64(w + 4 * (size_t)sib + 2), q3 = ld_dev64(w + 4 * (size_t)sib + 3);
        const float sx = __uint_as_float((uint32_t)q0), sy = __uint_as_float((uint32_t)(q0 >> 32));
        const float sz = __uint_as_float((uint32_t)q1);
        const float tx = __uint_as_float((uint32_t)q2), ty = __uint_as_float((uint32_t)(q2 >> 32));
        const float tz = __uint_as_float((uint32_t)q3);
        if (pa == cur) {   // this one is the left child
            lx = wmin(lx, sx); ly = wmin(ly, sy); lz = wmin(lz, sz);
            hx = wmax(hx, tx); hy = wmax(hy, ty); hz = wmax(hz, tz);
        } else {
            lx = wmin(sx, lx); ly = wmin(sy, ly); lz = wmin(sz, lz);
            hx = wmax(tx, hx); hy = wmax(ty, hy); hz = wmax(tz, hz);
        }
        put(p, pa, pb);
        cur = p;
    }
}

// the tree's cost (bvh_cost, engine/rt/bvh.c:464-469): the records' box areas over the root's, in
// double, one work-group
__global__ __launch_bounds__(1024) void k_cost(const float4 *__restrict__ nodes, int nn, double *out)
{
    __shared__ double s_sum[1024];
    double sum = 0.0;
    for (int i = threadIdx.x; i < nn; i += 1024) {
        const float4 lo = nodes[2 * (size_t)i], hi = nodes[2 * (size_t)i + 1];
        const double dx = (double)hi.x - lo.x, dy = (double)hi.y - lo.y, dz = (double)hi.z - lo.z;
        if (dx >= 0.0 && dy >= 0.0 && dz >= 0.0) sum += dx * dy + dy * dz + dz * dx;
    }
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float4 lo = nodes[0], hi = nodes[1];
        const double dx = (double)hi.x - lo.x, dy = (double)hi.y - lo.y, dz = (double)hi.z - lo.z;
        const double root = dx * dy + dy * dz + dz * dx;
        *out = root > 0.0 ? s_sum[0] / root : 0.0;
    }
}

static int top_cost(LrhPtScene *s, double *cost)
{
    LrhContext *c = s->ctx;
    hipLaunchKernelGGL(k_cost, dim3(1), dim3(1024), 0, c->stream, (const float4 *)s->tnodes, 2 * s->ninst - 1, s->tcost_dev);
    LR_HIP(hipGetLastError());
    LR_HIP(hipMemcpyAsync(cost, s->tcost_dev, 8, hipMemcpyDeviceToHost, c->stream));
    LR_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// The top level over the instances again, from their records on the device: their world boxes,
// the tree over them, its binary child-box records and its 4-wide copy; the walks' stack depth.
// Waits for it (collapse_wide does).
static int rebuild_top(LrhPtScene *s, int builder, LrhBuildStats *stats)
{
    LrhContext *c = s->ctx;
    const int ninst = s->ninst;
    hipLaunchKernelGGL(k_inst_boxes, dim3((unsigned)((ninst + 255) / 256)), dim3(256), 0, c->stream, s->ifwd, s->inst,
                       s->mroot, (const float4 *)s->mnodes, ninst, s->tboxes);
    LR_HIP(hipGetLastError());
    LrhBuildStats st;
    int rc = lrh_build(c, (const float *)s->tboxes, ninst, nullptr, builder, 0, 0, 0, s->tnodes, s->taux, nullptr, &st);
    if (rc) return rc;
    const size_t nn = 2 * (size_t)ninst - 1;
    hipLaunchKernelGGL(k_child_boxes, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, (const float4 *)s->tnodes, 0,
                       (int)nn, s->wnodes);
    if (ninst == 1) hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(1), 0, c->stream, (const float4 *)s->tnodes, 0u, s->wnodes);
    LR_HIP(hipGetLastError());
    // the top level 4-wide (the closest-hit walk's); collapse_wide waits for it
    rc = collapse_wide(c, s->tnodes, 0u, ninst, st.depth, s->tw4nodes);
    if (rc) return rc;
    // for later refits: each record's parent, and this build's cost
    hipLaunchKernelGGL(k_parents, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, (const float4 *)s->tnodes,
                       (int)nn, s->tparent);
    LR_HIP(hipGetLastError());
    rc = top_cost(s, &s->tcost_build);
    if (rc) return rc;
    s->tlas_depth = st.depth;
    s->top_builder = builder;
    // the two-level walks' stacks: per wide level of either tree at most three pending
    // entries (the binary walk: a sibling and an instance per top-level level), and EXIT
    s->depth = 3 * (s->tlas_depth + s->mesh_depth) + 8;
    st.refitted = 0;
    if (stats) *stats = st;
    return 0;
}

// The top level refitted to the instances' records (LRH_REFIT), kept while its cost stays within
// LR_REFIT_GROWTH of its last build's, else built again with that build's builder
static int refit_top(LrhPtScene *s, LrhBuildStats *stats)
{
    LrhContext *c = s->ctx;
    const int ninst = s->ninst, nn = 2 * ninst - 1;
    LR_HIP(hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_inst_boxes, dim3((unsigned)((ninst + 255) / 256)), dim3(256), 0, c->stream, s->ifwd, s->inst,
                       s->mroot, (const float4 *)s->mnodes, ninst, s->tboxes);
    hipLaunchKernelGGL(k_clear32, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, s->tarrive, nn);
    hipLaunchKernelGGL(k_refit, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, (float4 *)s->tnodes, nn, s->taux,
                       (const float4 *)s->tboxes, (const uint32_t *)s->tparent, s->tarrive);
    LR_HIP(hipGetLastError());
    double cost = 0.0;
    int rc = top_cost(s, &cost);
    if (rc) return rc;
    if (!(cost <= LR_REFIT_GROWTH * s->tcost_build)) return rebuild_top(s, s->top_builder, stats);
    hipLaunchKernelGGL(k_child_boxes, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, (const float4 *)s->tnodes, 0,
                       nn, s->wnodes);
    if (ninst == 1) hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(1), 0, c->stream, (const float4 *)s->tnodes, 0u, s->wnodes);
    LR_HIP(hipGetLastError());
    rc = collapse_wide(c, s->tnodes, 0u, ninst, s->tlas_depth, s->tw4nodes);   // (waits for it)
    if (rc) return rc;
    LR_HIP(hipEventRecord(c->ev[1], c->stream));
    LR_HIP(hipEventSynchronize(c->ev[1]));
    float ms = 0.0f;
    LR_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->depth = s->tlas_depth;
        stats->ms = ms;
        stats->refitted = 1;
    }
    return 0;
}

static int top_again(LrhPtScene *s, int builder, LrhBuildStats *stats)
{
    return builder == LRH_REFIT ? refit_top(s, stats) : rebuild_top(s, builder, stats);
}

extern "C" int lrh_pt_scene_update(LrhPtScene *s, const LrhInstance *inst, int ninst, int builder, LrhBuildStats *stats)
{
    if (!s || !s->two_level || !inst || ninst != s->ninst || (builder != LRH_LBVH && builder != LRH_PLOC && builder != LRH_REFIT))
        return lr_fail(LRH_EINVAL, "lrh_pt_scene_update: bad arguments (an instanced scene, the same instance count)");
    LrhContext *c = s->ctx;
    std::vector<float> rec, fwd;
    int rc = pack_instances(s, inst, ninst, s->mroot_h, s->mroot4_h, rec, fwd);
    if (rc) return rc;
    LR_HIP(hipStreamSynchronize(c->stream));
    LR_HIP(hipMemcpy(s->inst, rec.data(), rec.size() * 4, hipMemcpyHostToDevice));
    LR_HIP(hipMemcpy(s->ifwd, fwd.data(), fwd.size() * 4, hipMemcpyHostToDevice));
    s->hinst.assign(inst, inst + ninst);
    rc = emitters_again(s);
    if (rc) return rc;
    return top_again(s, builder, stats);
}

// moved instances' records, packed in one upload (index, 4 float4 record, 3 float4 rows each),
// into their places
__global__ __launch_bounds__(256) void k_scatter_inst(const int *__restrict__ idx, const float4 *__restrict__ rec,
                                                      const float4 *__restrict__ fwd, int n, float4 *__restrict__ inst,
                                                      float4 *__restrict__ ifwd)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const size_t i = (size_t)idx[k];
    for (int q = 0; q < 4; q++) inst[4 * i + q] = rec[4 * (size_t)k + q];
    for (int q = 0; q < 3; q++) ifwd[3 * i + q] = fwd[3 * (size_t)k + q];
}

extern "C" int lrh_pt_scene_move(LrhPtScene *s, const int *idx, const LrhInstance *inst, int n, int builder,
                                 LrhBuildStats *stats)
{
    if (!s || !s->two_level || n < 0 || (n > 0 && (!idx || !inst)) ||
        (builder != LRH_LBVH && builder != LRH_PLOC && builder != LRH_REFIT))
        return lr_fail(LRH_EINVAL, "lrh_pt_scene_move: bad arguments (an instanced scene, n >= 0)");
    std::vector<char> seen((size_t)s->ninst, 0);
    for (int k = 0; k < n; k++) {
        if (idx[k] < 0 || idx[k] >= s->ninst) return lr_fail(LRH_EINVAL, "lrh_pt_scene_move: instance %d of %d", idx[k], s->ninst);
        if (seen[(size_t)idx[k]]) return lr_fail(LRH_EINVAL, "lrh_pt_scene_move: instance %d given twice", idx[k]);
        seen[(size_t)idx[k]] = 1;
    }
    if (n == 0) {   // nothing moved: nothing sent, the top level stands
        if (stats) memset(stats, 0, sizeof(*stats));
        return 0;
    }
    std::vector<float> rec, fwd;
    int rc = pack_instances(s, inst, n, s->mroot_h, s->mroot4_h, rec, fwd);
    if (rc) return rc;
    const size_t bi = lr_align((size_t)n * 4), br = (size_t)n * 64, bf = (size_t)n * 48;
    std::vector<char> host(bi + br + bf);
    memcpy(host.data(), idx, (size_t)n * 4);
    memcpy(host.data() + bi, rec.data(), br);
    memcpy(host.data() + bi + br, fwd.data(), bf);
    LrhContext *c = s->ctx;
    LR_HIP(hipStreamSynchronize(c->stream));
    char *stage = nullptr;
    if (hipMalloc((void **)&stage, host.size()) != hipSuccess) {
        (void)hipGetLastError();
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene_move: device memory for %d records could not be had", n);
    }
    rc = 0;
    if (hipMemcpy(stage, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = lr_fail(LRH_EDEVICE, "lrh_pt_scene_move: the upload failed");
    if (!rc) {
        hipLaunchKernelGGL(k_scatter_inst, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (const int *)stage,
                           (const float4 *)(stage + bi), (const float4 *)(stage + bi + br), n, s->inst, s->ifwd);
        if (hipGetLastError() != hipSuccess) rc = lr_fail(LRH_EDEVICE, "lrh_pt_scene_move: a launch failed");
    }
    if (!rc) {
        for (int k = 0; k < n; k++) s->hinst[(size_t)idx[k]] = inst[k];
        rc = emitters_again(s);
    }
    if (!rc) rc = top_again(s, builder, stats);   // (waits for the stream)
    else (void)hipStreamSynchronize(c->stream);
    (void)hipFree(stage);
    return rc;
}

extern "C" int lrh_pt_scene_top(LrhPtScene *s, uint32_t *nodes, uint32_t *aux, double *cost)
{
    if (!s || !s->two_level) return lr_fail(LRH_EINVAL, "lrh_pt_scene_top: an instanced scene");
    LrhContext *c = s->ctx;
    const size_t nn = 2 * (size_t)s->ninst - 1;
    LR_HIP(hipStreamSynchronize(c->stream));
    if (nodes) LR_HIP(hipMemcpy(nodes, s->tnodes, nn * 32, hipMemcpyDeviceToHost));
    if (aux) LR_HIP(hipMemcpy(aux, s->taux, (size_t)s->ninst * 4, hipMemcpyDeviceToHost));
    if (cost) return top_cost(s, cost);
    return 0;
}

extern "C" int lrh_pt_scene_mesh(LrhPtScene *s, int mesh, const float *tris, int ntris, int builder, LrhBuildStats *stats)
{
    if (!s || !s->two_level || mesh < 0 || mesh >= s->nmeshes || !tris || (builder != LRH_LBVH && builder != LRH_PLOC))
        return lr_fail(LRH_EINVAL, "lrh_pt_scene_mesh: bad arguments (an instanced scene, one of its meshes)");
    const int f = s->mfirst[(size_t)mesh], n = s->mfirst[(size_t)mesh + 1] - f;
    if (ntris != n) return lr_fail(LRH_EINVAL, "lrh_pt_scene_mesh: mesh %d has %d triangles, not %d", mesh, n, ntris);
    for (int i = 0; i < n; i++) {
        uint32_t w;
        memcpy(&w, tris + 12 * (size_t)i + 11, 4);
        if ((int)(w >> 26) >= s->nmats) return lr_fail(LRH_EINVAL, "lrh_pt_scene_mesh: triangle %d: material %u", i, w >> 26);
    }
    LrhContext *c = s->ctx;
    const uint32_t root = s->mroot_h[(size_t)mesh];
    LR_HIP(hipStreamSynchronize(c->stream));
    LR_HIP(hipMemcpy(s->tris + 3 * (size_t)f, tris, (size_t)n * 48, hipMemcpyHostToDevice));
    // its boxes, its tree in its place (records from its root, aux entries from its first
    // triangle), the records' child boxes, its 4-wide copy from its first wide slot, its
    // triangles in leaf order
    int rc = lrh_tri_boxes(c, (const float *)(s->tris + 3 * (size_t)f), n, (float *)(s->mboxes + 2 * (size_t)f));
    LrhBuildStats st;
    if (!rc) rc = lrh_build(c, (const float *)(s->mboxes + 2 * (size_t)f), n, nullptr, builder, root, (uint32_t)f, (uint32_t)f,
                            s->mnodes, s->maux, nullptr, &st);
    if (rc) return rc;
    const int nn = 2 * n - 1;
    if (n == 1) hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(1), 0, c->stream, (const float4 *)s->mnodes, root, s->mwnodes);
    else
        hipLaunchKernelGGL(k_child_boxes, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, (const float4 *)s->mnodes,
                           (int)root, nn, s->mwnodes);
    LR_HIP(hipGetLastError());
    rc = collapse_wide(c, s->mnodes, root, n, st.depth, s->mw4nodes, (uint32_t)f);
    if (rc) return rc;
    hipLaunchKernelGGL(k_leaf_tris, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, s->maux + f, n, 0u,
                       (const float4 *)s->tris, s->ltris + 3 * (size_t)f);
    LR_HIP(hipGetLastError());
    s->mdepth[(size_t)mesh] = st.depth;
    s->mesh_depth = 0;
    for (int d : s->mdepth) s->mesh_depth = d > s->mesh_depth ? d : s->mesh_depth;
    // the instances' world boxes follow the mesh's: the top level again, with its own builder
    rc = rebuild_top(s, s->top_builder, nullptr);
    if (rc) return rc;
    if (stats) *stats = st;
    return 0;
}

extern "C" int lrh_pt_scene_instanced(LrhContext *c, const float *mtris, const int *mesh_first, int nmeshes,
                                      const LrhMaterial *mats, int nmats, const LrhInstance *inst, int ninst, float eps,
                                      int builder, LrhPtScene **out)
{
    if (!c || !out || nmeshes < 0 || !mesh_first || ninst <= 0 || !inst || !mats || nmats <= 0 || nmats > 64 ||
        (builder != LRH_LBVH && builder != LRH_PLOC))
        return lr_fail(LRH_EINVAL, "lrh_pt_scene_instanced: bad arguments");
    if (mesh_first[0] != 0) return lr_fail(LRH_EINVAL, "lrh_pt_scene_instanced: mesh_first[0] must be 0");
    for (int m = 0; m < nmeshes; m++)
        if (mesh_first[m + 1] <= mesh_first[m]) return lr_fail(LRH_EINVAL, "lrh_pt_scene_instanced: mesh %d is empty", m);
    const int M = mesh_first[nmeshes];
    if (M > 0 && !mtris) return lr_fail(LRH_EINVAL, "lrh_pt_scene_instanced: no triangles");
    if (M > LR_WIDE_MAX_SHAPES || ninst > LR_WIDE_MAX_SHAPES)
        return lr_fail(LRH_EREFUSED, "lrh_pt_scene_instanced: the wide nodes' leaf references take at most %d shapes",
                       LR_WIDE_MAX_SHAPES);
    for (int i = 0; i < M; i++) {
        uint32_t w;
        memcpy(&w, mtris + 12 * (size_t)i + 11, 4);
        if ((int)(w >> 26) >= nmats) return lr_fail(LRH_EINVAL, "lrh_pt_scene_instanced: triangle %d: material %u", i, w >> 26);
    }
    for (int i = 0; i < nmats; i++)
        if (mats[i].type < 0 || mats[i].type > LRH_MAT_PBR) return lr_fail(LRH_EINVAL, "material %d: type %d", i, mats[i].type);
    *out = nullptr;
    LrhPtScene *s = new LrhPtScene();
    s->ctx = c;
    s->two_level = 1;
    s->nmeshes = nmeshes;
    s->ninst = ninst;
    s->ntris = M;
    s->nmats = nmats;
    s->light = -1;
    s->eps = eps;
    // shadow rays to glowing shapes stop this far before them: ten times the rays' offset, as
    // RT_SHADOW_MARGIN is RT_EPS's in both of the reference's builds (tracer.c:15-23)
    s->margin = 10.0f * eps;
    s->hmats.assign(mats, mats + nmats);
    size_t nnodes = 0;
    std::vector<uint32_t> roots(nmeshes);
    for (int m = 0; m < nmeshes; m++) {
        roots[m] = (uint32_t)nnodes;
        nnodes += 2 * (size_t)(mesh_first[m + 1] - mesh_first[m]) - 1;
    }
    std::vector<float> hm(12 * (size_t)nmats);
    for (int i = 0; i < nmats; i++) {
        const LrhMaterial &m = mats[i];
        float *o = &hm[12 * (size_t)i];
        o[0] = m.albedo[0]; o[1] = m.albedo[1]; o[2] = m.albedo[2]; o[3] = (float)m.type;
        o[4] = m.emission[0]; o[5] = m.emission[1]; o[6] = m.emission[2]; o[7] = m.ior;
        o[8] = m.metallic; o[9] = m.roughness; o[10] = m.transmission; o[11] = 0.0f;
    }
    std::vector<float> tab(2112);
    std::vector<uint16_t> mask(4096);
    lrs_brdf_tables(tab.data());
    lrs_blue_noise(mask.data());
    const size_t tn = 2 * (size_t)ninst - 1;
    bool ok = dev_upload((void **)&s->tris, mtris, (size_t)M * 48) && dev_upload((void **)&s->mboxes, nullptr, (size_t)M * 32) &&
              dev_upload((void **)&s->mnodes, nullptr, nnodes * 32) && dev_upload((void **)&s->maux, nullptr, (size_t)M * 4) &&
              dev_upload((void **)&s->mwnodes, nullptr, nnodes * 64) &&
              dev_upload((void **)&s->mroot, roots.data(), roots.size() * 4) &&
              dev_upload((void **)&s->mats, hm.data(), hm.size() * 4) && dev_upload((void **)&s->tab, tab.data(), tab.size() * 4) &&
              dev_upload((void **)&s->mask, mask.data(), mask.size() * 2) &&
              dev_upload((void **)&s->inst, nullptr, (size_t)ninst * 64) && dev_upload((void **)&s->ifwd, nullptr, (size_t)ninst * 48) &&
              dev_upload((void **)&s->tboxes, nullptr, (size_t)ninst * 32) && dev_upload((void **)&s->tnodes, nullptr, tn * 32) &&
              dev_upload((void **)&s->taux, nullptr, (size_t)ninst * 4) && dev_upload((void **)&s->wnodes, nullptr, tn * 64) &&
              dev_upload((void **)&s->mw4nodes, nullptr, (size_t)M * 64) && dev_upload((void **)&s->ltris, nullptr, (size_t)M * 48) &&
              dev_upload((void **)&s->tw4nodes, nullptr, (size_t)ninst * 64) &&
              dev_upload((void **)&s->mroot4, nullptr, (size_t)nmeshes * 4) && dev_upload((void **)&s->tparent, nullptr, tn * 4) &&
              dev_upload((void **)&s->tarrive, nullptr, tn * 4) && dev_upload((void **)&s->tcost_dev, nullptr, 8) &&
              dev_upload((void **)&s->emit, nullptr, LR_MAX_EMITTERS * 80) &&
              dev_upload((void **)&s->inst_emit, nullptr, (size_t)ninst * 4);
    if (!ok) {
        (void)hipGetLastError();
        pt_scene_release(s);
        return lr_fail(LRH_ENOMEM, "lrh_pt_scene_instanced: device memory for the scene could not be had");
    }
    s->nodes = s->tnodes;
    s->aux = s->taux;
    s->boxes = s->tboxes;
    int rc = lrh_tri_boxes(c, mtris ? (const float *)s->tris : nullptr, M, (float *)s->mboxes);
    int depth = 0;
    std::vector<int> mdepth(nmeshes, 1);
    for (int m = 0; !rc && m < nmeshes; m++) {
        const int f = mesh_first[m], n = mesh_first[m + 1] - f;
        LrhBuildStats st;
        rc = lrh_build(c, (const float *)(s->mboxes + 2 * (size_t)f), n, nullptr, builder, roots[m], (uint32_t)f, (uint32_t)f,
                       s->mnodes, s->maux, nullptr, &st);
        if (!rc) mdepth[m] = st.depth;
        if (!rc && st.depth > depth) depth = st.depth;
        if (!rc && n == 1)
            hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(1), 0, c->stream, (const float4 *)s->mnodes, roots[m], s->mwnodes);
    }
    if (!rc && nnodes)
        hipLaunchKernelGGL(k_child_boxes, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, c->stream,
                           (const float4 *)s->mnodes, 0, (int)nnodes, s->mwnodes);
    // each mesh 4-wide, its wide nodes from the mesh's first triangle on (a mesh of n
    // triangles has at most n wide nodes), and the triangles in the meshes' leaf order
    std::vector<uint32_t> roots4(nmeshes);
    for (int m = 0; !rc && m < nmeshes; m++) {
        const int f = mesh_first[m], n = mesh_first[m + 1] - f;
        roots4[m] = (uint32_t)f;
        rc = collapse_wide(c, s->mnodes, roots[m], n, mdepth[m], s->mw4nodes, (uint32_t)f);
    }
    if (!rc && M) {
        hipLaunchKernelGGL(k_leaf_tris, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream, s->maux, M, 0u,
                           (const float4 *)s->tris, s->ltris);
        if (hipGetLastError() != hipSuccess) rc = lr_fail(LRH_EDEVICE, "lrh_pt_scene_instanced: a launch failed");
    }
    if (!rc && nmeshes && hipMemcpy(s->mroot4, roots4.data(), (size_t)nmeshes * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = lr_fail(LRH_EDEVICE, "lrh_pt_scene_instanced: a copy failed");
    if (rc) {
        pt_scene_release(s);
        return rc;
    }
    // the one-shape meshes' root records were written before the pass over all inner nodes,
    // which skips leaves: they stay
    s->mesh_depth = depth;
    s->mfirst.assign(mesh_first, mesh_first + nmeshes + 1);
    s->mdepth = mdepth;
    s->mroot_h = roots;
    s->mroot4_h = roots4;
    s->bytes = (size_t)M * (48 + 32 + 4 + 64 + 48) + nnodes * (32 + 64) + hm.size() * 4 + tab.size() * 4 + mask.size() * 2 +
               (size_t)ninst * (64 + 48 + 32 + 4 + 64 + 4) + tn * (32 + 64 + 8) + (size_t)nmeshes * 8 + 8 +
               LR_MAX_EMITTERS * 80;
    rc = lrh_pt_scene_update(s, inst, ninst, LRH_LBVH, nullptr);
    if (rc) {
        pt_scene_release(s);
        return rc;
    }
    *out = s;
    return 0;
}

void lr_job_scene(JobDev &jd, const LrhPtScene *s)
{
    jd.nodes = s->two_level ? s->tw4nodes : s->w4nodes;
    jd.bnodes = s->wnodes;
    jd.aux = s->aux;
    jd.tris = s->tris;
    jd.ltris = s->ltris;
    jd.prims = s->prims;
    jd.mats = s->mats;
    jd.root = 0;
    jd.nprims = (uint32_t)s->nprims;
    jd.light = s->light;
    jd.eps = s->eps;
    jd.margin = s->margin;
    jd.env_rgb = s->env_rgb;
    jd.env_marg = s->env_marg;
    jd.env_cond = s->env_cond;
    jd.env_prob = s->env_prob;
    jd.env_w = s->env_w;
    jd.env_h = s->env_h;
    jd.env_scale = s->env_scale;
    jd.env_rot = s->env_rot;
    jd.env_rot_c = cosf(s->env_rot);
    jd.env_rot_s = sinf(s->env_rot);
    jd.two_level = s->two_level;
    jd.mnodes = s->mwnodes;
    jd.mnodes4 = s->mw4nodes;
    jd.maux = s->maux;
    jd.inst = s->inst;
    jd.emit = s->emit;
    jd.inst_emit = s->inst_emit;
    jd.nemit = s->nemit;
}

// a batch of rays through a resident scene's own walk (the engine's closest-hit walk: walk4, or
// walk4_tl for an instanced scene); hits t, id, b1 (an instanced scene: the instance, as u32), b2
template <bool ANY>
__global__ __launch_bounds__(128) void k_scene_trace(JobDev J, const float4 *__restrict__ rays, int n, float4 *__restrict__ hits,
                                                     uint32_t *spill, int spill_per, uint32_t *overflow)
{
    __shared__ uint32_t stk[LR_LDS_STACK][128];
    const int i = blockIdx.x * 128 + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2 * (size_t)i], r1 = rays[2 * (size_t)i + 1];
    const WalkStack S{&stk[0][threadIdx.x], 128, spill + i, (size_t)n, LR_LDS_STACK + spill_per};
    WalkCnt wc;
    const WalkHit h = J.two_level ? walk4_tl<ANY, false, LR_LDS_STACK>(J, v3of(r0), v3of(r1), r0.w, r1.w, S, overflow, wc)
                                  : walk4<ANY, false, LR_LDS_STACK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, v3of(r0),
                                                                    v3of(r1), r0.w, r1.w, S, overflow, wc);
    hits[i] = make_float4(h.t, __uint_as_float(h.id), h.b1, h.b2);
}

extern "C" int lrh_pt_scene_trace(LrhContext *c, const LrhPtScene *s, const float *d_rays, int n, int any, float *d_hits)
{
    if (!c || !s || n < 0 || (n > 0 && (!d_rays || !d_hits))) return lr_fail(LRH_EINVAL, "lrh_pt_scene_trace: bad arguments");
    if (n == 0) return 0;
    const int spill_per = s->depth + 2 > LR_LDS_STACK ? s->depth + 2 - LR_LDS_STACK : 0;
    int rc = lr_scratch_reserve(c, LR_SCRATCH_TRACE, lr_align((size_t)spill_per * n * 4 + 4));
    if (rc) return rc;
    JobDev jd;
    memset(&jd, 0, sizeof(jd));
    lr_job_scene(jd, s);
    const dim3 grid((n + 127) / 128);
    uint32_t *spill = (uint32_t *)c->scratch[LR_SCRATCH_TRACE].base;
    if (any)
        hipLaunchKernelGGL(k_scene_trace<true>, grid, dim3(128), 0, c->stream, jd, (const float4 *)d_rays, n, (float4 *)d_hits,
                           spill, spill_per, c->flags);
    else
        hipLaunchKernelGGL(k_scene_trace<false>, grid, dim3(128), 0, c->stream, jd, (const float4 *)d_rays, n, (float4 *)d_hits,
                           spill, spill_per, c->flags);
    LR_HIP(hipGetLastError());
    return 0;
}
