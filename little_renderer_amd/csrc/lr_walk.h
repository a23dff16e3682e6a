// lr_walk.h - the BVH walk and the shape tests, shared by the trace kernels (lr_trace.hip)
// and the path tracer's extension and shadow kernels (lr_engine_k.h).
//
// The walk and its arithmetic are trace.wgsl traverse (:270-392), hit_tri (:159-173),
// hit_sphere / hit_quad (:130-157) and box_t (:175-184) with safe_inv (media.wgsl:45-48):
// the same f32 operations in the same order (the library is built without fused
// multiply-add), so hits equal the oracle's to the bit on the same tree. The stack differs
// from the reference's 40 entries that drop what does not fit: the first LR_LDS_STACK
// entries of each ray live in LDS (interleaved by lane: no bank conflicts) and deeper ones
// spill to device memory, sized from the tree's depth, so nothing is ever dropped. Sixteen
// LDS entries (8 KB for a work-group of 128) leave the walk kernels at their register-bound
// occupancy; the binary walk pushes only the farther child, so deep entries are rare.
#pragma once

#include "lr_device.h"

#define LR_LDS_STACK 16

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3of(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 vmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 vscale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ float safe_inv(float d) { return fabsf(d) > 1e-12f ? 1.0f / d : 1e30f; }

// box_t (trace.wgsl:175-184): the entry distance, -1 on a miss
__device__ __forceinline__ float box_t(float4 lo, float4 hi, V3 o, V3 inv, float tmin, float tmax)
{
    const float t0x = (lo.x - o.x) * inv.x, t0y = (lo.y - o.y) * inv.y, t0z = (lo.z - o.z) * inv.z;
    const float t1x = (hi.x - o.x) * inv.x, t1y = (hi.y - o.y) * inv.y, t1z = (hi.z - o.z) * inv.z;
    const float enter = wmax(wmax(wmin(t0x, t1x), wmin(t0y, t1y)), wmax(wmin(t0z, t1z), tmin));
    const float leave = wmin(wmin(wmax(t0x, t1x), wmax(t0y, t1y)), wmin(wmax(t0z, t1z), tmax));
    return leave < enter ? -1.0f : enter;
}

// hit_tri (trace.wgsl:159-173): t in (tmin, tmax) and the barycentrics
__device__ __forceinline__ bool hit_tri(float4 a, float4 b, float4 c, V3 o, V3 d, float tmin, float tmax, float &t_out,
                                        float &u_out, float &v_out)
{
    const V3 q = v3of(a), u = v3of(b), v = v3of(c);
    const V3 pvec = vcross(d, v);
    const float det = vdot(u, pvec);
    if (fabsf(det) < 1e-12f) return false;
    const float inv = 1.0f / det;
    const V3 tvec = vsub(o, q);
    const float b1 = vdot(tvec, pvec) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    const V3 qvec = vcross(tvec, u);
    const float b2 = vdot(d, qvec) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    const float t = vdot(v, qvec) * inv;
    if (t <= tmin || t >= tmax) return false;
    t_out = t;
    u_out = b1;
    v_out = b2;
    return true;
}

// a primitive's 96-byte record (trace.wgsl:3-10): c centre+radius, q corner (w: 0 sphere,
// 1 quad), u edge (w: material), v edge (w: area), n unit normal (w: plane_d), w = (u x v)/|u x v|^2
struct Prim {
    float4 c, q, u, v, n, w;
};

// hit_sphere (trace.wgsl:130-144)
__device__ __forceinline__ float hit_sphere(const Prim &p, V3 o, V3 d, float tmin, float tmax)
{
    const V3 oc = vsub(v3of(p.c), o);
    const float a = vdot(d, d), hh = vdot(d, oc), c = vdot(oc, oc) - p.c.w * p.c.w;
    const float disc = hh * hh - a * c;
    if (disc < 0.0f) return -1.0f;
    const float sq = sqrtf(disc);
    float t = (hh - sq) / a;
    if (t <= tmin || t >= tmax) {
        t = (hh + sq) / a;
        if (t <= tmin || t >= tmax) return -1.0f;
    }
    return t;
}

// hit_quad (trace.wgsl:146-157)
__device__ __forceinline__ float hit_quad(const Prim &p, V3 o, V3 d, float tmin, float tmax, float &a_out, float &b_out)
{
    const V3 n = v3of(p.n);
    const float denom = vdot(n, d);
    if (fabsf(denom) < 1e-12f) return -1.0f;
    const float t = (p.n.w - vdot(n, o)) / denom;
    if (t <= tmin || t >= tmax) return -1.0f;
    const V3 rel = vsub(vadd(o, vscale(d, t)), v3of(p.q));
    const float a = vdot(v3of(p.w), vcross(rel, v3of(p.v)));
    const float b = vdot(v3of(p.w), vcross(v3of(p.u), rel));
    if (a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f) return -1.0f;
    a_out = a;
    b_out = b;
    return t;
}

__device__ __forceinline__ Prim load_prim(const float4 *prims, uint32_t id)
{
    const float4 *p = prims + 6 * (size_t)id;
    return Prim{p[0], p[1], p[2], p[3], p[4], p[5]};
}

// hit_sphere / hit_quad on a record read as it is needed (the quad's edges only once its plane
// is hit): the same arithmetic, fewer registers held in the walks
__device__ __forceinline__ float hit_prim(const float4 *__restrict__ prims, uint32_t id, V3 o, V3 d, float tmin, float tmax,
                                          float &a_out, float &b_out)
{
    const size_t p = 6 * (size_t)id;
    const float4 q = ldg4(prims, p + 1);
    if (q.w < 0.5f) {
        Prim s;
        s.c = ldg4(prims, p);
        return hit_sphere(s, o, d, tmin, tmax);
    }
    const float4 nn = ldg4(prims, p + 4);
    const V3 n = v3of(nn);
    const float denom = vdot(n, d);
    if (fabsf(denom) < 1e-12f) return -1.0f;
    const float t = (nn.w - vdot(n, o)) / denom;
    if (t <= tmin || t >= tmax) return -1.0f;
    const float4 u = ldg4(prims, p + 2), v = ldg4(prims, p + 3), w = ldg4(prims, p + 5);
    const V3 rel = vsub(vadd(o, vscale(d, t)), v3of(q));
    const float a = vdot(v3of(w), vcross(rel, v3of(v)));
    const float b = vdot(v3of(w), vcross(v3of(u), rel));
    if (a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f) return -1.0f;
    a_out = a;
    b_out = b;
    return t;
}

struct WalkHit {
    float t;
    uint32_t id;
    float b1, b2;
};

// The stack of one ray: LDS slots stk[k * stride] for k < STK (the walk's template
// argument), then spill[(k - STK) * sstride] up to cap entries in all.
struct WalkStack {
    uint32_t *stk;
    int stride;
    uint32_t *spill;
    size_t sstride;
    int cap;
};

// entry sp of a walk's stack (LDS for the first STK, device memory past them). The LDS entry is
// read whatever sp is (at a clamped place) and the device one only past STK, as an atomic load
// (which the compiler does not merge with the other), so that the two stay an LDS read and a
// rare global one instead of becoming one flat read through a selected address (a vector
// memory instruction on every pop)
template <int STK>
__device__ __forceinline__ uint32_t stk_get(const WalkStack &S, int sp)
{
    uint32_t x = S.stk[(sp < STK ? sp : STK - 1) * S.stride];
    if (sp >= STK) x = __hip_atomic_load(&S.spill[(size_t)(sp - STK) * S.sstride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    return x;
}

// The closest hit in (tmin, tmax), or (ANY) the first one, over shape ids taken from the
// leaves' aux entries: ids below nprims are primitives, the rest triangles (id - nprims).
//
// The nodes are visited in the reference's order (pop a node; test its two children's boxes;
// push the farther hit child, then the nearer, and pop again), but the nearer child is not
// pushed and popped: it stays in registers as the next entry, with its two child fields
// (lo.w, hi.w), which its parent's box test has already loaded. Only a popped (farther) child
// fetches those two words of its record; the reference fetches the whole record of every node
// it pops. Same visits, same tests, same hits: one dependent load fewer on each step down.
// COUNT adds the node records fetched (the root's, each inner node's two children, each
// popped node's) and the shapes tested to cn, ct.
template <bool ANY, bool PRIMS, bool COUNT, int STK = LR_LDS_STACK>
__device__ __forceinline__ WalkHit walk(const float4 *__restrict__ nodes, uint32_t root, const uint32_t *__restrict__ aux,
                                        const float4 *__restrict__ tris, const float4 *__restrict__ prims, uint32_t nprims,
                                        V3 o, V3 d, float tmin, float tmax, const WalkStack &S, uint32_t *overflow,
                                        uint32_t &cn, uint32_t &ct)
{
    const V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    WalkHit best{tmax, LR_NONE, 0.0f, 0.0f};
    const float *words = reinterpret_cast<const float *>(nodes);
    // the entry being visited: its record's lo.w and hi.w (inner: the children; leaf: the
    // first aux entry, then LR_LEAF | count)
    uint32_t na = __float_as_uint(words[8 * (size_t)root + 3]), nb = __float_as_uint(words[8 * (size_t)root + 7]);
    if (COUNT) cn++;
    int sp = 0;
    for (;;) {
        while (nb & LR_LEAF) {
            const uint32_t cnt = nb & 0x7fffffffu;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t sid = aux[na + k];
                if (COUNT) ct++;
                if (PRIMS && sid < nprims) {
                    const Prim p = load_prim(prims, sid);
                    float a = 0.0f, b = 0.0f;
                    const float t = p.q.w < 0.5f ? hit_sphere(p, o, d, tmin, best.t) : hit_quad(p, o, d, tmin, best.t, a, b);
                    if (t > 0.0f) {
                        best = WalkHit{t, sid, a, b};
                        if (ANY) return best;
                    }
                    continue;
                }
                const uint32_t ti = PRIMS ? sid - nprims : sid;
                const float4 ta = tris[3 * (size_t)ti], tb = tris[3 * (size_t)ti + 1], tc = tris[3 * (size_t)ti + 2];
                float t, u, v;
                if (hit_tri(ta, tb, tc, o, d, tmin, best.t, t, u, v)) {
                    best = WalkHit{t, sid, u, v};
                    if (ANY) return best;
                }
            }
            if (sp == 0) return best;
            sp--;
            const uint32_t ni = stk_get<STK>(S, sp);
            na = __float_as_uint(words[8 * (size_t)ni + 3]);
            nb = __float_as_uint(words[8 * (size_t)ni + 7]);
            if (COUNT) cn++;
        }
        const float4 alo = nodes[2 * (size_t)na], ahi = nodes[2 * (size_t)na + 1];
        const float4 blo = nodes[2 * (size_t)nb], bhi = nodes[2 * (size_t)nb + 1];
        const float tl = box_t(alo, ahi, o, inv, tmin, best.t);
        const float tr = box_t(blo, bhi, o, inv, tmin, best.t);
        if (COUNT) cn += 2;
        if (tl >= 0.0f && tr >= 0.0f) {
            if (sp + 1 > S.cap) {   // cannot happen with a correct depth: said, never silently dropped
                atomicOr(overflow, 1u);
                break;
            }
            const bool left_first = tl <= tr;
            const uint32_t far = left_first ? nb : na;
            if (sp < STK) S.stk[sp * S.stride] = far;
            else S.spill[(size_t)(sp - STK) * S.sstride] = far;
            sp++;
            na = __float_as_uint(left_first ? alo.w : blo.w);
            nb = __float_as_uint(left_first ? ahi.w : bhi.w);
        } else if (tl >= 0.0f) {
            na = __float_as_uint(alo.w);
            nb = __float_as_uint(ahi.w);
        } else if (tr >= 0.0f) {
            na = __float_as_uint(blo.w);
            nb = __float_as_uint(bhi.w);
        } else {
            if (sp == 0) break;
            sp--;
            const uint32_t ni = stk_get<STK>(S, sp);
            na = __float_as_uint(words[8 * (size_t)ni + 3]);
            nb = __float_as_uint(words[8 * (size_t)ni + 7]);
            if (COUNT) cn++;
        }
    }
    return best;
}


// ---- the engine's walks: leaves -------------------------------------------------------------
// A walk's own counters (its counting variants): node records fetched and shapes tested by the
// lane, and the steps its wave took to do them (each counted once, by the wave's first lane at
// work in it), so that fetched / (64 x node steps) and tested / (64 x shape steps) are the shares
// of a wave's lanes at work in its node and shape steps.
struct WalkCnt {
    uint32_t n = 0, t = 0, wn = 0, wt = 0, wu = 0;   // wu: node steps whose lanes all fetch one node
};
__device__ __forceinline__ bool wave_first_lane() { return lane_id() == (uint32_t)__builtin_ctzll(__ballot(1)); }
__device__ __forceinline__ void count_node(WalkCnt &c)
{
    c.n++;
    if (wave_first_lane()) c.wn++;
}
// a node step at node k of the tree at w: counted as uniform when every lane at work in it is at
// the same node of the same tree
__device__ __forceinline__ void count_uniform(WalkCnt &c, const void *w, uint32_t k)
{
    const uint64_t p = (uint64_t)w;
    const uint32_t fk = __builtin_amdgcn_readfirstlane(k);
    const uint32_t flo = __builtin_amdgcn_readfirstlane((uint32_t)p), fhi = __builtin_amdgcn_readfirstlane((uint32_t)(p >> 32));
    const bool same = k == fk && (uint32_t)p == flo && (uint32_t)(p >> 32) == fhi;
    const uint64_t differ = __ballot(!same);
    if (wave_first_lane() && differ == 0) c.wu++;
}
__device__ __forceinline__ void count_shape(WalkCnt &c)
{
    c.t++;
    if (wave_first_lane()) c.wt++;
}

// ltris: the triangles in leaf order (aux entry e's triangle at ltris[3 e..]), so that the
// triangle's load does not wait for the aux entry's.
__device__ __forceinline__ void leaf_hits_(const uint32_t *__restrict__ aux, const float4 *__restrict__ ltris,
                                           const float4 *__restrict__ prims, uint32_t nprims, uint32_t first,
                                           uint32_t cnt, V3 o, V3 d, float tmin, WalkHit &best, bool any, bool &done,
                                           WalkCnt &c, bool count)
{
    for (uint32_t k = 0; k < cnt; k++) {
        const uint32_t e = first + k;
        const uint32_t sid = ldg(aux, e);
        const float4 ta = ldg4(ltris, 3 * (size_t)e), tb = ldg4(ltris, 3 * (size_t)e + 1), tc = ldg4(ltris, 3 * (size_t)e + 2);
        if (count) count_shape(c);
        if (sid < nprims) {
            float a = 0.0f, b = 0.0f;
            const float t = hit_prim(prims, sid, o, d, tmin, best.t, a, b);
            if (t > 0.0f) {
                best = WalkHit{t, sid, a, b};
                if (any) {
                    done = true;
                    return;
                }
            }
            continue;
        }
        float t, u, v;
        if (hit_tri(ta, tb, tc, o, d, tmin, best.t, t, u, v)) {
            best = WalkHit{t, sid, u, v};
            if (any) {
                done = true;
                return;
            }
        }
    }
}


// ---- the engine's 4-wide walk ---------------------------------------------------------------
// A wide node is 64 bytes (made by k_collapse, lr_pt_scene.hip, from the binary tree with the
// collapse rule of bvh.c:497-530): its box corner and per-axis power-of-two scales, the four
// children's references (a wide node's index; a leaf LR_LEAF | (count - 1) << 24 | first aux
// entry; none LR_NONE), and the children's boxes as 8-bit steps from the corner, rounded
// outwards.


// the entry distances (3e38: missed, or no child) and references of a wide node's children: the
// slab test with each plane's t as one fused multiply-add of its step, fma(q, s * inv, (corner -
// o) * inv) (a culling test on boxes that hold the children's: its last bit decides no hit).
// The step s * inv has the sign of inv (s > 0, or 0), and a fused multiply-add is monotone in q,
// so along an axis whose inv is negative the high plane is the near one: each axis's near and far
// planes are picked once a node, not ordered by a min and a max a child. A hit's entry is at most
// tmax (at most 1e30), so 3e38 sorts every miss last.
__device__ __forceinline__ void wide_test(const float4 *__restrict__ w, uint32_t node, V3 oi, V3 inv, float tmin, float tmax,
                                          float t[4], uint32_t ref[4])
{
    const float4 n0 = ldg4(w, 4 * (size_t)node), n1 = ldg4(w, 4 * (size_t)node + 1);
    const float4 n2 = ldg4(w, 4 * (size_t)node + 2), n3 = ldg4(w, 4 * (size_t)node + 3);
    const uint32_t eb = __float_as_uint(n0.w);
    const float ax = __uint_as_float((eb & 0xffu) << 23) * inv.x;
    const float ay = __uint_as_float(((eb >> 8) & 0xffu) << 23) * inv.y;
    const float az = __uint_as_float(((eb >> 16) & 0xffu) << 23) * inv.z;
    const float bx = __builtin_fmaf(n0.x, inv.x, -oi.x), by = __builtin_fmaf(n0.y, inv.y, -oi.y);
    const float bz = __builtin_fmaf(n0.z, inv.z, -oi.z);
    const uint32_t lx = __float_as_uint(n2.x), ly = __float_as_uint(n2.y), lz = __float_as_uint(n2.z);
    const uint32_t hx = __float_as_uint(n2.w), hy = __float_as_uint(n3.x), hz = __float_as_uint(n3.y);
    const bool fx = inv.x < 0.0f, fy = inv.y < 0.0f, fz = inv.z < 0.0f;   // (flipped: high plane near)
    const uint32_t nx = fx ? hx : lx, ny = fy ? hy : ly, nz = fz ? hz : lz;
    const uint32_t gx = fx ? lx : hx, gy = fy ? ly : hy, gz = fz ? lz : hz;
    ref[0] = __float_as_uint(n1.x);
    ref[1] = __float_as_uint(n1.y);
    ref[2] = __float_as_uint(n1.z);
    ref[3] = __float_as_uint(n1.w);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int sh = 8 * k;
        const float t0x = __builtin_fmaf((float)((nx >> sh) & 0xffu), ax, bx);
        const float t0y = __builtin_fmaf((float)((ny >> sh) & 0xffu), ay, by);
        const float t0z = __builtin_fmaf((float)((nz >> sh) & 0xffu), az, bz);
        const float t1x = __builtin_fmaf((float)((gx >> sh) & 0xffu), ax, bx);
        const float t1y = __builtin_fmaf((float)((gy >> sh) & 0xffu), ay, by);
        const float t1z = __builtin_fmaf((float)((gz >> sh) & 0xffu), az, bz);
        const float enter = fmaxf(fmaxf(t0x, t0y), fmaxf(t0z, tmin));
        const float leave = fminf(fminf(t1x, t1y), fminf(t1z, tmax));
        t[k] = ref[k] == LR_NONE || leave < enter ? 3e38f : enter;
        // one child after the other: interleaving the four holds all 24 planes at once
        __builtin_amdgcn_sched_barrier(0);
    }
}

__device__ __forceinline__ void cswap_(float &ta, uint32_t &ea, float &tb, uint32_t &eb)
{
    const bool s = tb < ta;
    const float t = s ? tb : ta;
    const uint32_t e = s ? eb : ea;
    tb = s ? ta : tb;
    eb = s ? ea : eb;
    ta = t;
    ea = e;
}

// One entry at a time: a wide node (its children's boxes tested, the children hit pushed
// farthest first, leaves among them) or a leaf (its shapes tested). Leaves and nodes are thus
// visited near to far, and the leaf test is one copy of its code, not one per child slot: fewer
// registers held, more waves in flight. The loop is written leaves-inside: a wave tests the
// leaves its lanes hold, then takes its next node step together (the other nesting, node
// steps inside until every lane holds a leaf, measured twice as slow).
template <bool ANY, bool COUNT, int STK>
__device__ __forceinline__ WalkHit walk4(const float4 *__restrict__ w4, const uint32_t *__restrict__ aux,
                                         const float4 *__restrict__ ltris, const float4 *__restrict__ prims,
                                         uint32_t nprims, V3 o, V3 d, float tmin, float tmax, const WalkStack &S,
                                         uint32_t *overflow, WalkCnt &c)
{
    const V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    WalkHit best{tmax, LR_NONE, 0.0f, 0.0f};
    int sp = 0;
    bool done = false;
    uint32_t cur = 0u;   // the entry being visited (the root node first); the next one stays in a register
    for (;;) {
        while (cur & LR_LEAF) {
            leaf_hits_(aux, ltris, prims, nprims, cur & 0xffffffu, ((cur >> 24) & 0x3fu) + 1u, o, d, tmin, best, ANY, done,
                       c, COUNT);
            if (ANY && done) return best;
            if (sp == 0) return best;
            sp--;
            cur = stk_get<STK>(S, sp);
        }
        if (COUNT) {
            count_node(c);
            count_uniform(c, w4, cur);
        }
        float t[4];
        uint32_t ref[4];
        // o * inv, the slab test's other term, made again each node step (fewer registers held)
        wide_test(w4, cur, v3(o.x * inv.x, o.y * inv.y, o.z * inv.z), inv, tmin, best.t, t, ref);
        // the children hit, nearest first: a sorting network over 4
        const int m = (t[0] < 3e38f) + (t[1] < 3e38f) + (t[2] < 3e38f) + (t[3] < 3e38f);
        float s0 = t[0], s1 = t[1], s2 = t[2], s3 = t[3];
        uint32_t c0 = ref[0], c1 = ref[1], c2 = ref[2], c3 = ref[3];
        cswap_(s0, c0, s1, c1);
        cswap_(s2, c2, s3, c3);
        cswap_(s0, c0, s2, c2);
        cswap_(s1, c1, s3, c3);
        cswap_(s1, c1, s2, c2);
        if (m == 0) {
            if (sp == 0) break;
            sp--;
            cur = stk_get<STK>(S, sp);
            continue;
        }
        if (sp + m - 1 > S.cap) {   // cannot happen with a correct depth: said, never silently dropped
            atomicOr(overflow, 1u);
            break;
        }
        auto push = [&](uint32_t x) {
            if (sp < STK) S.stk[sp * S.stride] = x;
            else S.spill[(size_t)(sp - STK) * S.sstride] = x;
            sp++;
        };
        if (m > 3) push(c3);
        if (m > 2) push(c2);
        if (m > 1) push(c1);
        cur = c0;
    }
    return best;
}
