// lr_rigid.hip - rigid bodies (spheres and boxes, in a room of six walls) stepped on the device:
// the rigid-body step that moves an instanced scene's dynamic instances between frames.
//
// The rules are the reference's CPU solver's (engine/phys/rigid.c): contacts found once per step
// with speculative margins (pair_margin, rigid.c:826-837) and followed through the substeps by
// their points in each body's frame; per substep gravity and damping, the contacts' rows, the
// warm start, sequential impulses with Coulomb friction, positions, and the relax pass that takes
// the push-out velocity back (rigid.c:1355-1545); a bounce once per contact from its approach
// speed when the step began. Sleeping and the swept test of fast bodies (rigid_ccd) are left out.
//
// The kernel is the device's own: one work-group of 1024 threads runs a whole step in one launch
// (no launch per phase or per iteration), its phases separated by barriers: the bodies' bounding
// spheres, every pair (i < j) tested in parallel and listed in (i, j) order by a scan, each item
// (a body's walls, a pair) colliding into slots of its own, the contacts compacted in the
// reference's order (a body's walls, then its pairs), the warm start matched against the last
// step's contacts. The contacts are coloured so that no two of one colour move the same body (in
// contact order, first free colour), and every solver pass runs colour after colour, the contacts
// of a colour in parallel: Gauss-Seidel like the reference's, in another order. Bodies and
// contacts live in device memory (L2-resident at these sizes); the colours' masks in LDS.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "lr_device.h"
#include "lr_internal.h"

namespace {

constexpr int RB_THREADS = 1024;
constexpr int RB_MAX_BODIES = 8192;        // the colour masks: two 64-bit words a body in LDS (128 KB)
constexpr int RB_COLOURS = 128;
constexpr int RB_WALL_SLOTS = 48;          // a box's corners against six walls
constexpr int RB_PAIR_SLOTS = 8;           // a clipped box face

// the reference's constants (rigid.c:13-28)
constexpr float CONTACT_SLOP = 0.5f;
constexpr float BAUMGARTE = 0.2f;
constexpr float BOUNCE_THRESHOLD = 60.0f;
constexpr float SPECULATIVE_MIN = 0.1f;
constexpr float MAX_PUSH = 300.0f;
constexpr float SLOW_SPEED = 60.0f;
constexpr float SETTLE_DAMPING = 2.0f;

// ---- small vector algebra ---------------------------------------------------------------------
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 f3of(float4 v) { return F3{v.x, v.y, v.z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 neg(F3 a) { return f3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b)
{
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len(F3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ float4 f4(F3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// a rotation: its matrix's columns (the body's axes in the world)
struct M3 {
    F3 c[3];
};
__device__ __forceinline__ M3 mat_of(float4 q)   // q = (x, y, z, w), a unit quaternion
{
    const float w = q.w, x = q.x, y = q.y, z = q.z;
    M3 r;
    r.c[0] = f3(1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y));
    r.c[1] = f3(2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x));
    r.c[2] = f3(2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y));
    return r;
}
__device__ __forceinline__ F3 rot(const M3 &r, F3 v) { return r.c[0] * v.x + r.c[1] * v.y + r.c[2] * v.z; }
__device__ __forceinline__ F3 rot_t(const M3 &r, F3 v) { return f3(dot(r.c[0], v), dot(r.c[1], v), dot(r.c[2], v)); }
// the world inverse inertia R diag(ii) R^T, symmetric: its six entries
struct S3 {
    float xx, yy, zz, xy, xz, yz;
};
__device__ __forceinline__ S3 inv_inertia_world(const M3 &r, F3 ii)
{
    S3 s;
    s.xx = ii.x * r.c[0].x * r.c[0].x + ii.y * r.c[1].x * r.c[1].x + ii.z * r.c[2].x * r.c[2].x;
    s.yy = ii.x * r.c[0].y * r.c[0].y + ii.y * r.c[1].y * r.c[1].y + ii.z * r.c[2].y * r.c[2].y;
    s.zz = ii.x * r.c[0].z * r.c[0].z + ii.y * r.c[1].z * r.c[1].z + ii.z * r.c[2].z * r.c[2].z;
    s.xy = ii.x * r.c[0].x * r.c[0].y + ii.y * r.c[1].x * r.c[1].y + ii.z * r.c[2].x * r.c[2].y;
    s.xz = ii.x * r.c[0].x * r.c[0].z + ii.y * r.c[1].x * r.c[1].z + ii.z * r.c[2].x * r.c[2].z;
    s.yz = ii.x * r.c[0].y * r.c[0].z + ii.y * r.c[1].y * r.c[1].z + ii.z * r.c[2].y * r.c[2].z;
    return s;
}
__device__ __forceinline__ F3 mul(const S3 &s, F3 v)
{
    return f3(s.xx * v.x + s.xy * v.y + s.xz * v.z, s.xy * v.x + s.yy * v.y + s.yz * v.z,
              s.xz * v.x + s.yz * v.y + s.zz * v.z);
}
// an orthonormal basis around a unit normal (Duff et al., "Building an orthonormal basis,
// revisited", JCGT 2017)
__device__ __forceinline__ void basis(F3 n, F3 &t, F3 &b)
{
    const float s = copysignf(1.0f, n.z);
    const float a = -1.0f / (s + n.z);
    const float c = n.x * n.y * a;
    t = f3(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    b = f3(c, s + n.y * n.y * a, -n.y);
}
// q advanced by the angular velocity w over h: q + h/2 (0, w) q, normalised
__device__ __forceinline__ float4 turn(float4 q, F3 w, float h)
{
    const float dw = -(w.x * q.x + w.y * q.y + w.z * q.z);
    const float dx = w.x * q.w + w.y * q.z - w.z * q.y;
    const float dy = w.y * q.w + w.z * q.x - w.x * q.z;
    const float dz = w.z * q.w + w.x * q.y - w.y * q.x;
    const float k = 0.5f * h;
    float4 r = make_float4(q.x + dx * k, q.y + dy * k, q.z + dz * k, q.w + dw * k);
    const float l = sqrtf(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
    return l > 0.0f ? make_float4(r.x / l, r.y / l, r.z / l, r.w / l) : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
}

}  // namespace

// ---- the world on the device ------------------------------------------------------------------
// a body: p (position, inverse mass), q (orientation x, y, z, w), v (velocity, restitution),
// w (angular velocity, friction), h (half sizes or radius, shape), ii (inverse inertia in the
// body's frame, mass)
struct RigidDev {
    float4 *p, *q, *v, *w, *h, *ii;
    float4 *iw;                  // per substep: the world inverse inertia (six entries, two float4)
    float4 *sph;                 // the bounding sphere grown by the step's motion (centre, radius)
    float *rad;                  // the bounding radius
    uint8_t *resting;            // touches something (settle damping)
    int *rowc;                   // pairs per body (row i: pairs (i, j > i)), then each row's first
    int2 *pairs;
    int *icount;                 // contacts per item (a body's walls, a pair), in item order
    float4 *slot;                // the items' contacts before compaction: point | normal, depth
    // contacts (this step's and the last one's: cur / prev, swapped by the host)
    struct Contacts {
        int2 *ab;                // bodies (b = -1: a wall)
        float4 *n;               // normal (from b to a), depth when found
        float4 *la, *lb;         // the point in a's frame, vn0 | in b's frame (a wall: the world), mu
        float4 *p;               // the point when found, on
        float4 *j;               // accumulated impulses jn, jt1, jt2; bias
        float4 *t1, *t2;         // tangents; mass_t1 | mass_t2 in w
        float4 *row;             // 9 float4 a contact: (ra x d, Ia^-1 (ra x d), rb x d / Ib^-1 (rb x d)) for n, t1, t2
        float4 *m;               // mass_n, depth now, bouncing, relax target
        int *colour;
    } c[2];
    int *order;                  // the contacts by colour
    int *cstart;                 // each colour's first place in order (RB_COLOURS + 1)
    int *prev_first;             // last step's first contact of each body (-1: none)
    int *counters;               // [0] contacts, [1] pairs, [2] colours, [3] overflow flags, [4] last step's contacts;
                                 // from [8]: the step's phases' device clock (5 x 64 bits: start, pairs, contacts,
                                 // colours, end)
};

struct RigidParamsDev {
    float gx, gy, gz;
    int has_room;
    float lo[3], hi[3];
    float wall_restitution, wall_friction;
    int iterations;
    float linear_damping, angular_damping;
    int warm_start;
};

namespace {

__device__ __forceinline__ float body_radius(float4 h) { return h.w < 0.5f ? h.x : sqrtf(h.x * h.x + h.y * h.y + h.z * h.z); }

// how far apart two bodies may be and still get a contact (pair_margin, rigid.c:826-837); b < 0: a wall
__device__ __forceinline__ float pair_margin(const RigidDev &D, int a, int b, float dt)
{
    const float ra = D.rad[a], rb = b >= 0 ? D.rad[b] : 0.0f;
    const F3 va = f3of(D.v[a]), wa = f3of(D.w[a]);
    float spin = len(wa) * ra, closing;
    if (b >= 0) {
        spin += len(f3of(D.w[b])) * rb;
        closing = len(va - f3of(D.v[b]));
    } else closing = len(va);
    return fminf(SPECULATIVE_MIN + dt * (closing + spin), SPECULATIVE_MIN + 0.5f * (ra + rb));
}

// a contact into an item's slots (point, normal from b to a, depth)
struct Sink {
    float4 *s;
    int n, cap;
    __device__ void add(F3 p, F3 nrm, float depth)
    {
        if (n < cap) {
            s[2 * n] = f4(p, 0.0f);
            s[2 * n + 1] = f4(nrm, depth);
        }
        n++;
    }
};

// the room's walls (collide_walls, rigid.c:470-491): planes dot(n, x) = off, n into the room
__device__ void collide_walls(const RigidDev &D, const RigidParamsDev &P, int i, float m, Sink &out)
{
    const F3 n[6] = {f3(1, 0, 0), f3(-1, 0, 0), f3(0, 1, 0), f3(0, -1, 0), f3(0, 0, 1), f3(0, 0, -1)};
    const float off[6] = {P.lo[0], -P.hi[0], P.lo[1], -P.hi[1], P.lo[2], -P.hi[2]};
    const F3 c = f3of(D.p[i]);
    const float4 h = D.h[i];
    if (h.w < 0.5f) {
        for (int k = 0; k < 6; k++) {
            const float d = dot(n[k], c) - off[k] - h.x;
            if (d < m) out.add(c - n[k] * h.x, n[k], -d);
        }
        return;
    }
    const M3 r = mat_of(D.q[i]);
    for (int k = 0; k < 6; k++)
        for (int q = 0; q < 8; q++) {
            const F3 l = f3((q & 1) ? h.x : -h.x, (q & 2) ? h.y : -h.y, (q & 4) ? h.z : -h.z);
            const F3 x = c + rot(r, l);
            const float d = dot(n[k], x) - off[k];
            if (d < m) out.add(x, n[k], -d);
        }
}

// (collide_sphere_sphere, rigid.c:633-643)
__device__ void sphere_sphere(const RigidDev &D, int a, int b, float m, Sink &out)
{
    const F3 d = f3of(D.p[a]) - f3of(D.p[b]);
    const float ra = D.h[a].x, rb = D.h[b].x, r = ra + rb, l2 = dot(d, d);
    if (l2 >= (r + m) * (r + m)) return;
    const float l = sqrtf(l2);
    const F3 n = l > 1e-6f ? d * (1.0f / l) : f3(0, 1, 0);
    const float depth = r - l;
    out.add(f3of(D.p[b]) + n * (rb - 0.5f * depth), n, depth);
}

// sphere s against box b, the normal from b to s (collide_sphere_box, rigid.c:645-685)
__device__ void sphere_box(const RigidDev &D, int s, int b, float m, Sink &out, bool flip)
{
    const M3 r = mat_of(D.q[b]);
    const F3 bp = f3of(D.p[b]), hb = f3of(D.h[b]);
    const F3 c = rot_t(r, f3of(D.p[s]) - bp);
    const F3 q = f3(fmaxf(-hb.x, fminf(hb.x, c.x)), fmaxf(-hb.y, fminf(hb.y, c.y)), fmaxf(-hb.z, fminf(hb.z, c.z)));
    const F3 diff = c - q;
    const float l2 = dot(diff, diff), rad = D.h[s].x;
    F3 nl;
    float depth;
    if (l2 > 1e-12f) {
        if (l2 >= (rad + m) * (rad + m)) return;
        const float l = sqrtf(l2);
        nl = diff * (1.0f / l);
        depth = rad - l;
    } else {   // the centre inside: out by the nearest face
        int k = 0;
        float best = 1e30f;
        for (int i = 0; i < 3; i++) {
            const float g = comp(hb, i) - fabsf(comp(c, i));
            if (g < best) {
                best = g;
                k = i;
            }
        }
        const float sg = comp(c, k) < 0.0f ? -1.0f : 1.0f;
        nl = f3(k == 0 ? sg : 0.0f, k == 1 ? sg : 0.0f, k == 2 ? sg : 0.0f);
        depth = rad + best;
    }
    const F3 n = rot(r, nl);
    const F3 p = bp + rot(r, q);
    out.add(p, flip ? neg(n) : n, depth);
}

// the part of polygon in (n points) where dot(x, d) <= off (Sutherland-Hodgman)
__device__ int clip(const F3 *in, int n, F3 d, float off, F3 *out)
{
    int k = 0;
    for (int i = 0; i < n; i++) {
        const F3 p = in[i], q = in[(i + 1) % n];
        const float dp = dot(p, d) - off, dq = dot(q, d) - off;
        if (dp <= 0.0f) out[k++] = p;
        if ((dp < 0.0f && dq > 0.0f) || (dp > 0.0f && dq < 0.0f)) out[k++] = p + (q - p) * (dp / (dp - dq));
    }
    return k;
}

// the closest points of segments p0 + s d0 and p1 + t d1 (s, t in [-1, 1])
__device__ void closest_segments(F3 p0, F3 d0, F3 p1, F3 d1, F3 &c0, F3 &c1)
{
    const F3 r = p0 - p1;
    const float a = dot(d0, d0), e = dot(d1, d1), f = dot(d1, r), c = dot(d0, r), b = dot(d0, d1);
    const float den = a * e - b * b;
    float s = 0.0f;
    if (den > 1e-8f) s = fmaxf(-1.0f, fminf(1.0f, (b * f - c * e) / den));
    float t = (b * s + f) / e;
    if (t < -1.0f || t > 1.0f) {
        t = fmaxf(-1.0f, fminf(1.0f, t));
        s = fmaxf(-1.0f, fminf(1.0f, (t * b - c) / a));
    }
    c0 = p0 + d0 * s;
    c1 = p1 + d1 * t;
}

// Two boxes (collide_box_box, rigid.c:724-816): the separating axis of least overlap among the
// 15 (face axes preferred, a's over b's); on a face axis the incident face of the other box is
// clipped to the reference face's sides and its points below the face are the contacts, else one
// contact between the two edges.
__device__ void box_box(const RigidDev &D, int a, int b, float m, Sink &out)
{
    const M3 ra = mat_of(D.q[a]), rb = mat_of(D.q[b]);
    const F3 pa = f3of(D.p[a]), pb = f3of(D.p[b]), ha = f3of(D.h[a]), hb = f3of(D.h[b]);
    const F3 t = pa - pb;
    float best = 1e30f, depth = 0.0f;
    F3 bn = f3(0, 1, 0);
    int best_axis = -1;
    for (int i = 0; i < 15; i++) {
        F3 L = i < 3 ? ra.c[i] : (i < 6 ? rb.c[i - 3] : cross(ra.c[(i - 6) / 3], rb.c[(i - 6) % 3]));
        const float ll = dot(L, L);
        if (ll < 1e-6f) continue;
        L = L * (1.0f / sqrtf(ll));
        const float qa = ha.x * fabsf(dot(ra.c[0], L)) + ha.y * fabsf(dot(ra.c[1], L)) + ha.z * fabsf(dot(ra.c[2], L));
        const float qb = hb.x * fabsf(dot(rb.c[0], L)) + hb.y * fabsf(dot(rb.c[1], L)) + hb.z * fabsf(dot(rb.c[2], L));
        const float dist = dot(t, L);
        const float overlap = qa + qb - fabsf(dist);
        if (overlap < -m) return;
        const float score = i < 3 ? overlap : (i < 6 ? overlap * 1.02f + 0.01f : overlap * 1.1f + 0.1f);
        if (score < best) {
            best = score;
            depth = overlap;
            bn = dist < 0.0f ? neg(L) : L;
            best_axis = i;
        }
    }
    if (best_axis < 0) return;
    if (best_axis < 6) {
        const bool ref_a = best_axis < 3;
        const M3 &rr = ref_a ? ra : rb, &ri = ref_a ? rb : ra;
        const F3 rp = ref_a ? pa : pb, ip = ref_a ? pb : pa, rh = ref_a ? ha : hb, ih = ref_a ? hb : ha;
        const F3 nref = ref_a ? neg(bn) : bn;
        const int k = best_axis % 3;
        int ki = 0;
        float most = 0.0f;
        for (int q = 0; q < 3; q++) {
            const float d = dot(ri.c[q], nref);
            if (fabsf(d) > fabsf(most)) {
                most = d;
                ki = q;
            }
        }
        // the incident face (axis ki, the side facing the reference box)
        F3 poly[8], tmp[8];   // (a quad clipped by four planes: at most 8 corners)
        {
            const float s = most > 0.0f ? -1.0f : 1.0f;
            const int i1 = (ki + 1) % 3, i2 = (ki + 2) % 3;
            const F3 c = ip + ri.c[ki] * (s * comp(ih, ki));
            const F3 u = ri.c[i1] * comp(ih, i1), v = ri.c[i2] * comp(ih, i2);
            poly[0] = c + u + v;
            poly[1] = c + (u - v);
            poly[2] = c - (u + v);
            poly[3] = c - (u - v);
        }
        int n = 4;
        for (int q = 1; q <= 2 && n > 0; q++) {
            const int j = (k + q) % 3;
            const F3 dj = rr.c[j];
            const float hj = comp(rh, j), cj = dot(dj, rp);
            n = clip(poly, n, dj, cj + hj, tmp);
            n = clip(tmp, n, neg(dj), -(cj - hj), poly);
        }
        const float face = dot(nref, rp) + comp(rh, k);
        int added = 0;
        for (int q = 0; q < n; q++) {
            const float sep = dot(nref, poly[q]) - face;
            if (sep > m) continue;
            out.add(poly[q] - nref * (0.5f * sep), bn, -sep);
            added++;
        }
        if (added) return;
    }
    const int ea = best_axis >= 6 ? (best_axis - 6) / 3 : 0, eb = best_axis >= 6 ? (best_axis - 6) % 3 : 0;
    F3 qa = pa, qb = pb;
    for (int q = 0; q < 3; q++) {
        if (q != ea) qa = qa + ra.c[q] * ((dot(ra.c[q], bn) > 0.0f ? -1.0f : 1.0f) * comp(ha, q));
        if (q != eb) qb = qb + rb.c[q] * ((dot(rb.c[q], bn) > 0.0f ? 1.0f : -1.0f) * comp(hb, q));
    }
    F3 c0, c1;
    closest_segments(qa, ra.c[ea] * comp(ha, ea), qb, rb.c[eb] * comp(hb, eb), c0, c1);
    out.add((c0 + c1) * 0.5f, bn, depth);
}

// exclusive scan over the work-group of one value a thread; *total its sum
__device__ int block_scan(int v, int *total, int *s_w)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < RB_THREADS / 64; k++) {
        const int s = s_w[k];
        if (k < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// exclusive scan of a[0..n) in place (in order), over the work-group; returns the total
__device__ int scan_array(int *a, int n, int *s_w)
{
    const int per = (n + RB_THREADS - 1) / RB_THREADS;
    const int i0 = threadIdx.x * per;
    int sum = 0;
    for (int k = 0; k < per && i0 + k < n; k++) sum += a[i0 + k];
    int total;
    int run = block_scan(sum, &total, s_w);
    for (int k = 0; k < per && i0 + k < n; k++) {
        const int x = a[i0 + k];
        a[i0 + k] = run;
        run += x;
    }
    __syncthreads();
    return total;
}

// one contact's rows along d (solve_row_make, rigid.c:1153-1158)
__device__ __forceinline__ void row_make(float4 *row, F3 d, F3 ra, F3 rb, const S3 &ia, const S3 *ib)
{
    const F3 rad = cross(ra, d), iad = mul(ia, rad);
    F3 rbd = f3(0, 0, 0), ibd = f3(0, 0, 0);
    if (ib) {
        rbd = cross(rb, d);
        ibd = mul(*ib, rbd);
    }
    row[0] = f4(rad, rbd.x);
    row[1] = f4(iad, rbd.y);
    row[2] = make_float4(rbd.z, ibd.x, ibd.y, ibd.z);
}

struct Row {
    F3 d, rad, iad, rbd, ibd;
};
__device__ __forceinline__ Row row_get(const float4 *row, F3 d)
{
    const float4 a = row[0], b = row[1], c = row[2];
    return Row{d, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(a.w, b.w, c.x), f3(c.y, c.z, c.w)};
}
__device__ __forceinline__ float row_speed(const RigidDev &D, int a, int b, const Row &r)
{
    float v = dot(r.d, f3of(D.v[a])) + dot(r.rad, f3of(D.w[a]));
    if (b >= 0) v -= dot(r.d, f3of(D.v[b])) + dot(r.rbd, f3of(D.w[b]));
    return v;
}
__device__ __forceinline__ void add3(float4 *x, F3 d)
{
    float4 v = *x;
    v.x += d.x;
    v.y += d.y;
    v.z += d.z;
    *x = v;
}
__device__ __forceinline__ void row_apply(const RigidDev &D, int a, int b, const Row &r, float j)
{
    const float ma = D.p[a].w;
    if (ma > 0.0f) {
        add3(&D.v[a], r.d * (j * ma));
        add3(&D.w[a], r.iad * j);
    }
    if (b >= 0) {
        const float mb = D.p[b].w;
        if (mb > 0.0f) {
            add3(&D.v[b], r.d * (-j * mb));
            add3(&D.w[b], r.ibd * (-j));
        }
    }
}

// one contact towards its target speed along the normal, friction after it (solve_contact,
// rigid.c:1178-1197)
__device__ void solve_contact(const RigidDev &D, const RigidDev::Contacts &C, int k, float bias)
{
    const int2 ab = C.ab[k];
    const float4 nd = C.n[k], t1 = C.t1[k], t2 = C.t2[k], mm = C.m[k];
    float4 j = C.j[k];
    const Row rn = row_get(C.row + 9 * (size_t)k, f3of(nd));
    float dj = mm.x * (bias - row_speed(D, ab.x, ab.y, rn));
    const float jn = fmaxf(j.x + dj, 0.0f);
    dj = jn - j.x;
    j.x = jn;
    if (dj != 0.0f) row_apply(D, ab.x, ab.y, rn, dj);
    if (j.x != 0.0f || j.y != 0.0f || j.z != 0.0f) {
        const float lim = C.lb[k].w * j.x;
        const Row r1 = row_get(C.row + 9 * (size_t)k + 3, f3of(t1));
        float jt = fmaxf(-lim, fminf(lim, j.y - t1.w * row_speed(D, ab.x, ab.y, r1)));
        row_apply(D, ab.x, ab.y, r1, jt - j.y);
        j.y = jt;
        const Row r2 = row_get(C.row + 9 * (size_t)k + 6, f3of(t2));
        jt = fmaxf(-lim, fminf(lim, j.z - t2.w * row_speed(D, ab.x, ab.y, r2)));
        row_apply(D, ab.x, ab.y, r2, jt - j.z);
        j.z = jt;
    }
    C.j[k] = j;
}

__device__ __forceinline__ float eff_mass(float ma, float mb, F3 rad, F3 iad, F3 rbd, F3 ibd)
{
    // 1 / (ma + mb + (Ia (ra x d)) . (ra x d) + (Ib (rb x d)) . (rb x d))
    const float k = ma + mb + dot(iad, rad) + dot(ibd, rbd);
    return k > 0.0f ? 1.0f / k : 0.0f;
}

// ---- the broad phase over the whole chip (broad_phase and test_pair, rigid.c:826-900) ----
// 1. each body's bounding sphere grown by its motion in the step; the step's per-body state reset;
// the step's first clock
constexpr int RB_WIDE = 256;
__global__ __launch_bounds__(RB_WIDE) void k_rigid_prep(RigidDev D, int nb, float dt)
{
    const int i = blockIdx.x * RB_WIDE + threadIdx.x;
    if (i == 0) reinterpret_cast<unsigned long long *>(D.counters + 8)[0] = wall_clock64();
    if (i >= nb) return;
    const float4 h = D.h[i];
    const float r = body_radius(h);
    D.rad[i] = r;
    const float grow = 0.5f * SPECULATIVE_MIN + dt * (len(f3of(D.v[i])) + len(f3of(D.w[i])) * r);
    D.sph[i] = f4(f3of(D.p[i]), r + grow);
    D.resting[i] = 0;
    D.prev_first[i] = -1;
}

// whether bodies i < j make a pair: their grown spheres meet and their margin lets them touch
// (test_pair); two fixed bodies never do
__device__ __forceinline__ bool rigid_pair(const RigidDev &D, int i, int j, float4 si, bool fi, float dt)
{
    if (fi && D.p[j].w == 0.0f) return false;
    const float4 sj = D.sph[j];
    const F3 d = f3(si.x - sj.x, si.y - sj.y, si.z - sj.z);
    const float rr = (si.w + sj.w) * 1.0001f + 1e-4f;
    if (dot(d, d) > rr * rr) return false;
    const float r = D.rad[i] + D.rad[j] + pair_margin(D, i, j, dt);
    return dot(d, d) <= r * r;
}

// 2. (EMIT false) each row's pairs counted into rowc[i]; (EMIT true) written from the row's first
// place, in j order: one wave a row, 64 j at a time, each lane's place from the ballot
template <bool EMIT>
__global__ __launch_bounds__(RB_WIDE) void k_rigid_rows(RigidDev D, int nb, float dt, int max_pairs)
{
    const int i = blockIdx.x * (RB_WIDE / 64) + (int)(threadIdx.x / 64);
    if (i >= nb) return;
    const int lane = (int)(threadIdx.x & 63);
    const float4 si = D.sph[i];
    const bool fi = D.p[i].w == 0.0f;
    int at = EMIT ? D.rowc[i] : 0;
    for (int j0 = i + 1; j0 < nb; j0 += 64) {
        const int j = j0 + lane;
        const bool hit = j < nb && rigid_pair(D, i, j, si, fi, dt);
        const uint64_t b = __ballot(hit);
        if (EMIT && hit) {
            const int q = at + __popcll(b & lanemask_lt());
            if (q < max_pairs) D.pairs[q] = make_int2(i, j);
        }
        at += __popcll(b);
    }
    if (!EMIT && lane == 0) D.rowc[i] = at;
}

// 3. the rows' counts to their first places (in order), the total to counters[5]
__global__ __launch_bounds__(RB_THREADS) void k_rigid_scan(RigidDev D, int nb)
{
    __shared__ int s_w[RB_THREADS / 64];
    const int np = scan_array(D.rowc, nb, s_w);
    if (threadIdx.x == 0) D.counters[5] = np;
}

// The step (rigid_steps, rigid.c:1371-1545), one work-group. cur / prev: this step's contacts
// and the last one's.
__global__ __launch_bounds__(RB_THREADS) void k_rigid_step(RigidDev D, RigidParamsDev P, int nb, int cur, float dt, int substeps,
                                                           int max_pairs, int max_contacts, int first_step)
{
    __shared__ int s_w[RB_THREADS / 64];
    __shared__ int s_total;
    __shared__ uint64_t mask[RB_MAX_BODIES][2];
    __shared__ int s_cc[RB_COLOURS];
    const int tid = threadIdx.x;
    const RigidDev::Contacts &C = D.c[cur], &Q = D.c[1 - cur];
    const int nprev = first_step ? 0 : D.counters[4];
    unsigned long long *clk = reinterpret_cast<unsigned long long *>(D.counters + 8);
    // the pairs: listed by the broad-phase kernels before it, rowc each row's first place
    const int np = D.counters[5];
    if (tid == 0) {
        D.counters[3] = 0;
        clk[1] = wall_clock64();
    }
    if (np > max_pairs) {
        if (tid == 0) {
            D.counters[1] = np;
            D.counters[3] |= 1;
        }
        return;
    }
    // ---- the items' contacts into their slots: item i + rowc[i] is body i's walls, item i + 1 +
    // p pair p (of row i): the reference's order (contacts_of, rigid.c:929-946) ----
    const int nitems = nb + np;
    for (int it = tid; it < nitems; it += RB_THREADS) {
        // which body's walls or which pair: the row whose walls item is the last at or before it
        int lo = 0, hi = nb - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (mid + D.rowc[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        const int i = lo, wall_item = i + D.rowc[i];
        Sink s;
        if (it == wall_item) {
            s = Sink{D.slot + 2 * (size_t)RB_WALL_SLOTS * i, 0, RB_WALL_SLOTS};
            if (P.has_room && D.p[i].w > 0.0f) collide_walls(D, P, i, pair_margin(D, i, -1, dt), s);
        } else {
            const int p = it - i - 1;
            const int2 pr = D.pairs[p];
            s = Sink{D.slot + 2 * (size_t)RB_WALL_SLOTS * nb + 2 * (size_t)RB_PAIR_SLOTS * p, 0, RB_PAIR_SLOTS};
            const float m = pair_margin(D, pr.x, pr.y, dt);
            const bool sa = D.h[pr.x].w < 0.5f, sb = D.h[pr.y].w < 0.5f;
            if (sa && sb) sphere_sphere(D, pr.x, pr.y, m, s);
            else if (!sa && !sb) box_box(D, pr.x, pr.y, m, s);
            else if (sa) sphere_box(D, pr.x, pr.y, m, s, false);
            else sphere_box(D, pr.y, pr.x, m, s, true);
        }
        D.icount[it] = s.n < s.cap ? s.n : s.cap;
        if (s.n > s.cap) atomicOr(&D.counters[3], 4);
    }
    __syncthreads();
    const int nc = scan_array(D.icount, nitems, s_w);   // icount: each item's first contact
    if (nc > max_contacts) {
        if (tid == 0) {
            D.counters[0] = nc;
            D.counters[3] |= 2;
        }
        return;
    }
    // ---- the contacts in order: bodies, normal, depth, the point in each body's frame
    // (add_contact, rigid.c:411-421), tangents; the warm start from the last step's same contact
    // (same bodies, within 1 unit, normals within 0.95: rigid.c:1393-1410) ----
    for (int q = tid; q < nprev; q += RB_THREADS) {
        const int a = Q.ab[q].x;
        if (q == 0 || Q.ab[q - 1].x != a) D.prev_first[a] = q;
    }
    __syncthreads();
    for (int it = tid; it < nitems; it += RB_THREADS) {
        int lo = 0, hi = nb - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (mid + D.rowc[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        const int i = lo;
        int a, b;
        const float4 *src;
        if (it == i + D.rowc[i]) {
            a = i;
            b = -1;
            src = D.slot + 2 * (size_t)RB_WALL_SLOTS * i;
        } else {
            const int p = it - i - 1;
            a = D.pairs[p].x;
            b = D.pairs[p].y;
            src = D.slot + 2 * (size_t)RB_WALL_SLOTS * nb + 2 * (size_t)RB_PAIR_SLOTS * p;
        }
        const int first = D.icount[it], cnt = (it + 1 < nitems ? D.icount[it + 1] : nc) - first;
        const M3 ra = mat_of(D.q[a]);
        const F3 pa = f3of(D.p[a]);
        M3 rbm;
        F3 pb = f3(0, 0, 0);
        if (b >= 0) {
            rbm = mat_of(D.q[b]);
            pb = f3of(D.p[b]);
        }
        for (int q = 0; q < cnt; q++) {
            const int k = first + q;
            const F3 x = f3of(src[2 * q]);
            const float4 nd = src[2 * q + 1];
            const F3 n = f3of(nd);
            C.ab[k] = make_int2(a, b);
            C.n[k] = nd;
            C.p[k] = f4(x, 0.0f);
            F3 t1, t2;
            basis(n, t1, t2);
            C.t1[k] = f4(t1, 0.0f);
            C.t2[k] = f4(t2, 0.0f);
            float4 j = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float vn0 = 0.0f;
            if (P.warm_start) {
                const int q0 = D.prev_first[a];
                if (q0 >= 0)
                    for (int o = q0; o < nprev && Q.ab[o].x == a; o++) {
                        if (Q.ab[o].y != b) continue;
                        const F3 d = f3of(Q.p[o]) - x;
                        if (dot(d, d) > 1.0f || dot(f3of(Q.n[o]), n) < 0.95f) continue;
                        const float4 oj = Q.j[o];
                        j = make_float4(oj.x, oj.y, oj.z, 0.0f);
                        if (oj.x > 0.0f) vn0 = Q.la[o].w;
                        break;
                    }
            }
            C.j[k] = j;
            C.la[k] = f4(rot_t(ra, x - pa), vn0);
            C.lb[k] = f4(b >= 0 ? rot_t(rbm, x - pb) : x, 0.0f);
            if (nd.w > -SPECULATIVE_MIN) {   // touching: its bodies settle (rigid.c:1414-1421)
                D.resting[a] = 1;
                if (b >= 0) D.resting[b] = 1;
            }
        }
    }
    __syncthreads();
    if (tid == 0) clk[2] = wall_clock64();
    // ---- colours: contact k the first colour neither of its moving bodies has yet, in order ----
    for (int i = tid; i < nb; i += RB_THREADS) mask[i][0] = mask[i][1] = 0ull;
    for (int k = tid; k < RB_COLOURS; k += RB_THREADS) s_cc[k] = 0;
    __syncthreads();
    // (by the first wave: its lanes read 64 contacts' bodies at once, lane 0 colours them in
    // order from the masks, and each lane keeps its contact's colour)
    if (tid < 64) {
        int ncol = 0;
        for (int base = 0; base < nc; base += 64) {
            const int k = base + tid;
            uint32_t word = 0xffffffffu;   // the moving bodies: a | b << 16, 0xffff none
            if (k < nc) {
                const int2 ab = C.ab[k];
                const uint32_t a = D.p[ab.x].w > 0.0f ? (uint32_t)ab.x : 0xffffu;
                const uint32_t b = ab.y >= 0 && D.p[ab.y].w > 0.0f ? (uint32_t)ab.y : 0xffffu;
                word = a | b << 16;
            }
            const int cnt = nc - base < 64 ? nc - base : 64;
            int mine = -1;
            for (int q = 0; q < cnt; q++) {
                const uint32_t wq = __shfl(word, q, 64);
                int col = -1;
                if (tid == 0) {
                    const uint32_t a = wq & 0xffffu, b = wq >> 16;
                    uint64_t m0 = 0, m1 = 0;
                    if (a != 0xffffu) {
                        m0 |= mask[a][0];
                        m1 |= mask[a][1];
                    }
                    if (b != 0xffffu) {
                        m0 |= mask[b][0];
                        m1 |= mask[b][1];
                    }
                    col = ~m0 ? __builtin_ctzll(~m0) : (~m1 ? 64 + __builtin_ctzll(~m1) : -1);
                    if (col >= 0) {
                        const uint64_t bit = 1ull << (col & 63);
                        if (a != 0xffffu) mask[a][col >> 6] |= bit;
                        if (b != 0xffffu) mask[b][col >> 6] |= bit;
                        s_cc[col]++;
                        if (col + 1 > ncol) ncol = col + 1;
                    } else D.counters[3] |= 8;   // more than RB_COLOURS on one body: not solved this step (said)
                }
                col = __shfl(col, 0, 64);
                if (tid == q) mine = col;
            }
            if (k < nc) C.colour[k] = mine;
        }
        if (tid == 0) {
            int run = 0;
            for (int c = 0; c < RB_COLOURS; c++) {
                D.cstart[c] = run;
                const int x = s_cc[c];
                s_cc[c] = run;   // (now each colour's next place)
                run += x;
            }
            D.cstart[RB_COLOURS] = run;
            s_total = ncol;
            D.counters[0] = nc;
            D.counters[1] = np;
            D.counters[2] = ncol;
        }
        // the contacts by colour, each colour's in contact order
        for (int base = 0; base < nc; base += 64) {
            const int k = base + tid;
            const int col = k < nc ? C.colour[k] : -1;
            const int cnt = nc - base < 64 ? nc - base : 64;
            int at = -1;
            for (int q = 0; q < cnt; q++) {
                const int cq = __shfl(col, q, 64);
                int pos = -1;
                if (tid == 0 && cq >= 0) pos = s_cc[cq]++;
                pos = __shfl(pos, 0, 64);
                if (tid == q) at = pos;
            }
            if (at >= 0) D.order[at] = k;
        }
    }
    __syncthreads();
    const int ncol = s_total;
    if (tid == 0) clk[3] = wall_clock64();
    // every pass over the contacts: colour after colour, a colour's contacts in parallel
    auto by_colour = [&](auto &&f) {
        for (int c = 0; c < ncol; c++) {
            const int e = D.cstart[c + 1];
            for (int q = D.cstart[c] + tid; q < e; q += RB_THREADS) f(D.order[q]);
            __syncthreads();
        }
    };
    const float h = dt / (float)substeps;
    const F3 g = f3(P.gx, P.gy, P.gz);
    for (int s = 0; s < substeps; s++) {
        // 1. gravity -> velocities, damping (a slow body resting on something settles), the
        // world inverse inertia for this substep
        for (int i = tid; i < nb; i += RB_THREADS) {
            const float4 p = D.p[i];
            const M3 r = mat_of(D.q[i]);
            const S3 iw = p.w > 0.0f ? inv_inertia_world(r, f3of(D.ii[i])) : S3{0, 0, 0, 0, 0, 0};
            D.iw[2 * i] = make_float4(iw.xx, iw.yy, iw.zz, iw.xy);
            D.iw[2 * i + 1] = make_float4(iw.xz, iw.yz, 0.0f, 0.0f);
            if (p.w == 0.0f) continue;
            float4 v = D.v[i], w = D.w[i];
            F3 vv = f3of(v) + g * h, ww = f3of(w);
            const bool slow = len(vv) < SLOW_SPEED && len(ww) * D.rad[i] < SLOW_SPEED;
            const float settle = D.resting[i] && slow ? SETTLE_DAMPING : 0.0f;
            vv = vv * (1.0f / (1.0f + (P.linear_damping + settle) * h));
            ww = ww * (1.0f / (1.0f + (P.angular_damping + settle) * h));
            D.v[i] = f4(vv, v.w);
            D.w[i] = f4(ww, w.w);
        }
        __syncthreads();
        // 2. each contact where its bodies are now: arms, depth, effective masses, rows, target
        for (int k = tid; k < nc; k += RB_THREADS) {
            const int2 ab = C.ab[k];
            const int a = ab.x, b = ab.y;
            const float4 nd = C.n[k], la = C.la[k], lb = C.lb[k], t1 = C.t1[k], t2 = C.t2[k];
            const F3 n = f3of(nd);
            const M3 rma = mat_of(D.q[a]);
            const F3 ra = rot(rma, f3of(la));
            F3 rb = f3(0, 0, 0), pb = f3of(lb);
            if (b >= 0) {
                rb = rot(mat_of(D.q[b]), f3of(lb));
                pb = f3of(D.p[b]) + rb;
            }
            const F3 pa = f3of(D.p[a]) + ra;
            const float depth = nd.w - dot(pa - pb, n);
            const float4 i0 = D.iw[2 * a], i1 = D.iw[2 * a + 1];
            const S3 ia{i0.x, i0.y, i0.z, i0.w, i1.x, i1.y};
            S3 ib{0, 0, 0, 0, 0, 0};
            if (b >= 0) {
                const float4 j0 = D.iw[2 * b], j1 = D.iw[2 * b + 1];
                ib = S3{j0.x, j0.y, j0.z, j0.w, j1.x, j1.y};
            }
            float4 *row = C.row + 9 * (size_t)k;
            row_make(row, n, ra, rb, ia, b >= 0 ? &ib : nullptr);
            row_make(row + 3, f3of(t1), ra, rb, ia, b >= 0 ? &ib : nullptr);
            row_make(row + 6, f3of(t2), ra, rb, ia, b >= 0 ? &ib : nullptr);
            const float ma = D.p[a].w, mb = b >= 0 ? D.p[b].w : 0.0f;
            const Row rn = row_get(row, n), r1 = row_get(row + 3, f3of(t1)), r2 = row_get(row + 6, f3of(t2));
            const float mass_n = eff_mass(ma, mb, rn.rad, rn.iad, rn.rbd, rn.ibd);
            C.t1[k] = f4(f3of(t1), eff_mass(ma, mb, r1.rad, r1.iad, r1.rbd, r1.ibd));
            C.t2[k] = f4(f3of(t2), eff_mass(ma, mb, r2.rad, r2.iad, r2.rbd, r2.ibd));
            const float vn = row_speed(D, a, b, rn);
            float vn0 = la.w;
            if (s == 0) vn0 = fminf(vn0, vn);
            const float ea = D.v[a].w, fa = D.w[a].w;
            const float e = b >= 0 ? fmaxf(ea, D.v[b].w) : fmaxf(ea, P.wall_restitution);
            const float mu = b >= 0 ? sqrtf(fa * D.w[b].w) : sqrtf(fa * P.wall_friction);
            const float bounce = depth > -SPECULATIVE_MIN && vn0 < -BOUNCE_THRESHOLD ? -e * vn0 : 0.0f;
            float4 j = C.j[k];
            j.w = bounce > 0.0f ? bounce
                                : (depth < 0.0f ? depth / h : fminf(BAUMGARTE / h * fmaxf(depth - CONTACT_SLOP, 0.0f), MAX_PUSH));
            bool on;
            if (bounce > 0.0f) {
                j.x = j.y = j.z = 0.0f;
                vn0 = 0.0f;
                on = true;
            } else on = depth > -SPECULATIVE_MIN || j.x > 0.0f || -depth < 2.0f * h * fmaxf(-vn, 0.0f) + SPECULATIVE_MIN;
            C.j[k] = j;
            C.la[k] = f4(f3of(la), vn0);
            C.lb[k] = f4(f3of(lb), mu);
            C.m[k] = make_float4(mass_n, depth, bounce > 0.0f ? 1.0f : 0.0f, 0.0f);
            C.p[k].w = on ? 1.0f : 0.0f;
        }
        __syncthreads();
        // the impulse each contact had (last step's, or the last substep's), applied first
        by_colour([&](int k) {
            const float4 j = C.j[k];
            if (C.m[k].z != 0.0f || (j.x == 0.0f && j.y == 0.0f && j.z == 0.0f)) return;
            const int2 ab = C.ab[k];
            const Row rn = row_get(C.row + 9 * (size_t)k, f3of(C.n[k])), r1 = row_get(C.row + 9 * (size_t)k + 3, f3of(C.t1[k])),
                      r2 = row_get(C.row + 9 * (size_t)k + 6, f3of(C.t2[k]));
            row_apply(D, ab.x, ab.y, rn, j.x);
            row_apply(D, ab.x, ab.y, r1, j.y);
            row_apply(D, ab.x, ab.y, r2, j.z);
        });
        // 3. the iterations
        for (int it = 0; it < P.iterations; it++)
            by_colour([&](int k) {
                if (C.p[k].w != 0.0f) solve_contact(D, C, k, C.j[k].w);
            });
        // 4. velocities -> positions
        for (int i = tid; i < nb; i += RB_THREADS) {
            const float4 p = D.p[i];
            if (p.w == 0.0f) continue;
            D.p[i] = f4(f3of(p) + f3of(D.v[i]) * h, p.w);
            D.q[i] = turn(D.q[i], f3of(D.w[i]), h);
        }
        __syncthreads();
        // 5. relax: towards the speeds the contacts allow without a push (rigid.c:1520-1533)
        for (int k = tid; k < nc; k += RB_THREADS) {
            if (C.p[k].w == 0.0f) continue;
            const int2 ab = C.ab[k];
            const float4 la = C.la[k], lb = C.lb[k];
            const F3 pa = f3of(D.p[ab.x]) + rot(mat_of(D.q[ab.x]), f3of(la));
            F3 pb = f3of(lb);
            if (ab.y >= 0) pb = f3of(D.p[ab.y]) + rot(mat_of(D.q[ab.y]), f3of(lb));
            const float depth = C.n[k].w - dot(pa - pb, f3of(C.n[k]));
            float4 m = C.m[k];
            m.w = m.z != 0.0f ? C.j[k].w : (depth < 0.0f ? depth / h : 0.0f);
            C.m[k] = m;
        }
        __syncthreads();
        by_colour([&](int k) {
            if (C.p[k].w != 0.0f) solve_contact(D, C, k, C.m[k].w);
        });
    }
    if (tid == 0) {
        D.counters[4] = nc;
        clk[4] = wall_clock64();
    }
}

}  // namespace

// ---- host side --------------------------------------------------------------------------------
struct LrhRigid {
    LrhContext *ctx = nullptr;
    int cap = 0, max_pairs = 0, max_contacts = 0;
    int n = 0;
    int cur = 0;
    bool stepped = false;       // a step has run: the device bodies are newer than the host's
    bool dirty = true;          // the host's bodies are newer than the device's
    std::vector<LrhBody> host;
    LrhRigidParams params{};
    RigidDev D{};
    void *mem = nullptr;
    LrhRigidStats stats{};
};

static int pull(LrhRigid *r);

namespace {

void body_fill(LrhBody &b, int shape, const float *c, const float *half, float density)
{
    memset(&b, 0, sizeof(b));
    b.shape = shape;
    for (int k = 0; k < 3; k++) {
        b.pos[k] = c[k];
        b.half[k] = half[k];
    }
    b.rot[3] = 1.0f;
    b.restitution = 0.4f;   // body_base (rigid.c:161-175)
    b.friction = 0.5f;
}

}  // namespace

extern "C" int lrh_rigid_create(LrhContext *c, int max_bodies, int max_contacts, LrhRigid **out)
{
    if (!c || !out || max_bodies < 1 || max_bodies > RB_MAX_BODIES || max_contacts < 1)
        return lr_fail(LRH_EINVAL, "lrh_rigid_create: bad arguments (1 <= max_bodies <= %d)", RB_MAX_BODIES);
    LrhRigid *r = new LrhRigid();
    r->ctx = c;
    r->cap = max_bodies;
    r->max_contacts = max_contacts;
    r->max_pairs = max_contacts;   // a pair listed gives a contact or none: as many as the contacts
    // rigid_init (rigid.c:122-132), without sleeping
    LrhRigidParams &p = r->params;
    p.gravity[1] = -981.0f;
    p.wall_restitution = 0.4f;
    p.wall_friction = 0.5f;
    p.iterations = 10;
    p.linear_damping = 0.0f;
    p.angular_damping = 0.1f;
    p.warm_start = 1;
    const size_t B = (size_t)max_bodies, P = (size_t)r->max_pairs, K = (size_t)max_contacts;
    std::vector<size_t> sz = {B * 16 * 6, B * 32, B * 16, B * 4, B, B * 4 + 4, P * 8, (B + P) * 4 + 4,
                              (2 * RB_WALL_SLOTS * B + 2 * RB_PAIR_SLOTS * P) * 16};
    for (int h = 0; h < 2; h++) {
        std::vector<size_t> cs = {K * 8, K * 16, K * 16, K * 16, K * 16, K * 16, K * 16, K * 16, K * 16 * 9, K * 16, K * 4};
        sz.insert(sz.end(), cs.begin(), cs.end());
    }
    const std::vector<size_t> tail = {K * 4, (RB_COLOURS + 1) * 4, B * 4, 128};
    sz.insert(sz.end(), tail.begin(), tail.end());
    size_t total = 0;
    for (size_t z : sz) total += lr_align(z);
    if (hipMalloc(&r->mem, total) != hipSuccess) {
        (void)hipGetLastError();
        delete r;
        return lr_fail(LRH_ENOMEM, "lrh_rigid_create: %zu bytes of device memory could not be had", total);
    }
    if (hipMemset(r->mem, 0, total) != hipSuccess) {
        (void)hipFree(r->mem);
        delete r;
        return lr_fail(LRH_EDEVICE, "lrh_rigid_create: clearing its memory failed");
    }
    char *at = (char *)r->mem;
    size_t q = 0;
    auto take = [&]() {
        char *x = at;
        at += lr_align(sz[q++]);
        return x;
    };
    RigidDev &D = r->D;
    float4 *bodies = (float4 *)take();
    D.p = bodies;
    D.q = bodies + B;
    D.v = bodies + 2 * B;
    D.w = bodies + 3 * B;
    D.h = bodies + 4 * B;
    D.ii = bodies + 5 * B;
    D.iw = (float4 *)take();
    D.sph = (float4 *)take();
    D.rad = (float *)take();
    D.resting = (uint8_t *)take();
    D.rowc = (int *)take();
    D.pairs = (int2 *)take();
    D.icount = (int *)take();
    D.slot = (float4 *)take();
    for (int h = 0; h < 2; h++) {
        RigidDev::Contacts &C = D.c[h];
        C.ab = (int2 *)take();
        C.n = (float4 *)take();
        C.la = (float4 *)take();
        C.lb = (float4 *)take();
        C.p = (float4 *)take();
        C.j = (float4 *)take();
        C.t1 = (float4 *)take();
        C.t2 = (float4 *)take();
        C.row = (float4 *)take();
        C.m = (float4 *)take();
        C.colour = (int *)take();
    }
    D.order = (int *)take();
    D.cstart = (int *)take();
    D.prev_first = (int *)take();
    D.counters = (int *)take();
    *out = r;
    return 0;
}

extern "C" void lrh_rigid_destroy(LrhRigid *r)
{
    if (!r) return;
    (void)hipStreamSynchronize(r->ctx->stream);
    (void)hipFree(r->mem);
    delete r;
}

extern "C" int lrh_rigid_params_get(const LrhRigid *r, LrhRigidParams *p)
{
    if (!r || !p) return lr_fail(LRH_EINVAL, "lrh_rigid_params_get: bad arguments");
    *p = r->params;
    return 0;
}

extern "C" int lrh_rigid_params_set(LrhRigid *r, const LrhRigidParams *p)
{
    if (!r || !p || p->iterations < 0) return lr_fail(LRH_EINVAL, "lrh_rigid_params_set: bad arguments");
    if (p->has_room && !(p->room_lo[0] < p->room_hi[0] && p->room_lo[1] < p->room_hi[1] && p->room_lo[2] < p->room_hi[2]))
        return lr_fail(LRH_EINVAL, "lrh_rigid_params_set: an empty room");
    r->params = *p;
    return 0;
}

// rigid_add_sphere / rigid_add_box (rigid.c:177-201): mass from the density, inertia of a
// solid ball / box, restitution 0.4, friction 0.5
extern "C" int lrh_rigid_add_sphere(LrhRigid *r, const float *c, float radius, float density)
{
    if (!r || !c || !(radius > 0.0f) || !(density >= 0.0f)) return lr_fail(LRH_EINVAL, "lrh_rigid_add_sphere: bad arguments");
    if (r->n >= r->cap) return lr_fail(LRH_EREFUSED, "lrh_rigid_add_sphere: the world holds %d bodies", r->cap);
    int rc = pull(r);
    if (rc) return rc;
    LrhBody b;
    const float half[3] = {radius, radius, radius};
    body_fill(b, LRH_BODY_SPHERE, c, half, density);
    const float mass = 4.0f / 3.0f * 3.14159265358979f * radius * radius * radius * density;
    b.inv_mass = mass > 0.0f ? 1.0f / mass : 0.0f;
    const float i = mass > 0.0f ? 1.0f / (0.4f * mass * radius * radius) : 0.0f;
    b.inv_inertia[0] = b.inv_inertia[1] = b.inv_inertia[2] = i;
    r->host.push_back(b);
    r->dirty = true;
    return r->n++;
}

extern "C" int lrh_rigid_add_box(LrhRigid *r, const float *c, const float *half, const float *q, float density)
{
    if (!r || !c || !half || !q || !(half[0] > 0.0f && half[1] > 0.0f && half[2] > 0.0f) || !(density >= 0.0f))
        return lr_fail(LRH_EINVAL, "lrh_rigid_add_box: bad arguments");
    if (r->n >= r->cap) return lr_fail(LRH_EREFUSED, "lrh_rigid_add_box: the world holds %d bodies", r->cap);
    int rc = pull(r);
    if (rc) return rc;
    LrhBody b;
    body_fill(b, LRH_BODY_BOX, c, half, density);
    const float l = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; k++) b.rot[k] = l > 0.0f ? q[k] / l : (k == 3 ? 1.0f : 0.0f);
    const float mass = 8.0f * half[0] * half[1] * half[2] * density;
    if (mass > 0.0f) {
        b.inv_mass = 1.0f / mass;
        b.inv_inertia[0] = 3.0f / (mass * (half[1] * half[1] + half[2] * half[2]));
        b.inv_inertia[1] = 3.0f / (mass * (half[0] * half[0] + half[2] * half[2]));
        b.inv_inertia[2] = 3.0f / (mass * (half[0] * half[0] + half[1] * half[1]));
    }
    r->host.push_back(b);
    r->dirty = true;
    return r->n++;
}

// the device's bodies into the host's copy (after steps)
static int pull(LrhRigid *r)
{
    if (!r->stepped || r->n == 0) return 0;
    const size_t B = (size_t)r->cap, n = (size_t)r->n;
    std::vector<float4> x(6 * n);
    LrhContext *c = r->ctx;
    for (int f = 0; f < 4; f++)
        LR_HIP(hipMemcpyAsync(x.data() + f * n, r->D.p + f * B, n * 16, hipMemcpyDeviceToHost, c->stream));
    LR_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) {
        LrhBody &b = r->host[i];
        const float4 p = x[i], q = x[n + i], v = x[2 * n + i], w = x[3 * n + i];
        b.pos[0] = p.x; b.pos[1] = p.y; b.pos[2] = p.z;
        b.rot[0] = q.x; b.rot[1] = q.y; b.rot[2] = q.z; b.rot[3] = q.w;
        b.vel[0] = v.x; b.vel[1] = v.y; b.vel[2] = v.z;
        b.angvel[0] = w.x; b.angvel[1] = w.y; b.angvel[2] = w.z;
    }
    r->stepped = false;
    return 0;
}

// the host's bodies onto the device (before a step, when changed)
static int push(LrhRigid *r)
{
    if (!r->dirty) return 0;
    const size_t B = (size_t)r->cap, n = (size_t)r->n;
    std::vector<float4> x(6 * n);
    for (size_t i = 0; i < n; i++) {
        const LrhBody &b = r->host[i];
        x[i] = make_float4(b.pos[0], b.pos[1], b.pos[2], b.inv_mass);
        x[n + i] = make_float4(b.rot[0], b.rot[1], b.rot[2], b.rot[3]);
        x[2 * n + i] = make_float4(b.vel[0], b.vel[1], b.vel[2], b.restitution);
        x[3 * n + i] = make_float4(b.angvel[0], b.angvel[1], b.angvel[2], b.friction);
        x[4 * n + i] = make_float4(b.half[0], b.half[1], b.half[2], b.shape == LRH_BODY_BOX ? 1.0f : 0.0f);
        const float mass = b.inv_mass > 0.0f ? 1.0f / b.inv_mass : 0.0f;
        x[5 * n + i] = make_float4(b.inv_inertia[0], b.inv_inertia[1], b.inv_inertia[2], mass);
    }
    LrhContext *c = r->ctx;
    if (n)
        for (int f = 0; f < 6; f++)
            LR_HIP(hipMemcpyAsync(r->D.p + f * B, x.data() + f * n, n * 16, hipMemcpyHostToDevice, c->stream));
    LR_HIP(hipStreamSynchronize(c->stream));
    r->dirty = false;
    return 0;
}

extern "C" int lrh_rigid_get(LrhRigid *r, int first, int n, LrhBody *out)
{
    if (!r || first < 0 || n < 0 || first + n > r->n || (n > 0 && !out)) return lr_fail(LRH_EINVAL, "lrh_rigid_get: bad range");
    int rc = pull(r);
    if (rc) return rc;
    if (n) memcpy(out, r->host.data() + first, (size_t)n * sizeof(LrhBody));
    return 0;
}

extern "C" int lrh_rigid_set(LrhRigid *r, int first, int n, const LrhBody *in)
{
    if (!r || first < 0 || n < 0 || first + n > r->n || (n > 0 && !in)) return lr_fail(LRH_EINVAL, "lrh_rigid_set: bad range");
    for (int k = 0; k < n; k++)
        if ((in[k].shape != LRH_BODY_SPHERE && in[k].shape != LRH_BODY_BOX) || !(in[k].half[0] > 0.0f) || in[k].inv_mass < 0.0f)
            return lr_fail(LRH_EINVAL, "lrh_rigid_set: body %d: a sphere or a box, sizes > 0, inverse mass >= 0", first + k);
    int rc = pull(r);
    if (rc) return rc;
    if (n) memcpy(r->host.data() + first, in, (size_t)n * sizeof(LrhBody));
    r->dirty = true;
    return 0;
}

extern "C" int lrh_rigid_steps(LrhRigid *r, float dt, int substeps, int nsteps)
{
    if (!r || !(dt > 0.0f) || substeps < 1 || nsteps < 0) return lr_fail(LRH_EINVAL, "lrh_rigid_steps: bad arguments");
    if (r->n == 0 || nsteps == 0) return 0;
    int rc = pull(r);   // (the host's copy is current before it is sent)
    if (rc) return rc;
    rc = push(r);
    if (rc) return rc;
    const LrhRigidParams &p = r->params;
    RigidParamsDev P;
    P.gx = p.gravity[0];
    P.gy = p.gravity[1];
    P.gz = p.gravity[2];
    P.has_room = p.has_room;
    for (int k = 0; k < 3; k++) {
        P.lo[k] = p.room_lo[k];
        P.hi[k] = p.room_hi[k];
    }
    P.wall_restitution = p.wall_restitution;
    P.wall_friction = p.wall_friction;
    P.iterations = p.iterations;
    P.linear_damping = p.linear_damping;
    P.angular_damping = p.angular_damping;
    P.warm_start = p.warm_start;
    LrhContext *c = r->ctx;
    LR_HIP(hipEventRecord(c->ev[0], c->stream));
    for (int s = 0; s < nsteps; s++) {
        r->cur = 1 - r->cur;
        const int nb = r->n;
        hipLaunchKernelGGL(k_rigid_prep, dim3((nb + RB_WIDE - 1) / RB_WIDE), dim3(RB_WIDE), 0, c->stream, r->D, nb, dt);
        hipLaunchKernelGGL(k_rigid_rows<false>, dim3((nb + RB_WIDE / 64 - 1) / (RB_WIDE / 64)), dim3(RB_WIDE), 0, c->stream,
                           r->D, nb, dt, r->max_pairs);
        hipLaunchKernelGGL(k_rigid_scan, dim3(1), dim3(RB_THREADS), 0, c->stream, r->D, nb);
        hipLaunchKernelGGL(k_rigid_rows<true>, dim3((nb + RB_WIDE / 64 - 1) / (RB_WIDE / 64)), dim3(RB_WIDE), 0, c->stream,
                           r->D, nb, dt, r->max_pairs);
        hipLaunchKernelGGL(k_rigid_step, dim3(1), dim3(RB_THREADS), 0, c->stream, r->D, P, nb, r->cur, dt, substeps,
                           r->max_pairs, r->max_contacts, r->stats.steps == 0 && s == 0 ? 1 : 0);
        LR_HIP(hipGetLastError());
    }
    LR_HIP(hipEventRecord(c->ev[1], c->stream));
    LR_HIP(hipEventSynchronize(c->ev[1]));
    float ms = 0.0f;
    LR_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    int cnt[18];
    LR_HIP(hipMemcpy(cnt, r->D.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    unsigned long long clk[5];
    memcpy(clk, cnt + 8, sizeof clk);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) {
        (void)hipGetLastError();
        khz = 100000;
    }
    for (int k = 0; k < 4; k++) r->stats.phase_ms[k] = (float)((double)(clk[k + 1] - clk[k]) / (double)khz);
    r->stepped = true;
    r->stats.steps += nsteps;
    r->stats.contacts = cnt[0];
    r->stats.pairs = cnt[1];
    r->stats.colours = cnt[2];
    r->stats.ms = ms / (float)nsteps;
    if (cnt[3] & 1) return lr_fail(LRH_EREFUSED, "lrh_rigid_steps: %d pairs passed the world's room for %d", cnt[1], r->max_pairs);
    if (cnt[3] & 2)
        return lr_fail(LRH_EREFUSED, "lrh_rigid_steps: %d contacts passed the world's room for %d", cnt[0], r->max_contacts);
    if (cnt[3] & 12) return lr_fail(LRH_EDEVICE, "lrh_rigid_steps: a pair's contacts or a body's colours ran out (flags %d)", cnt[3]);
    return 0;
}

extern "C" int lrh_rigid_stats(const LrhRigid *r, LrhRigidStats *st)
{
    if (!r || !st) return lr_fail(LRH_EINVAL, "lrh_rigid_stats: bad arguments");
    *st = r->stats;
    return 0;
}
