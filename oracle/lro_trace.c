/* lro_trace.c - the oracle's ray traversals: the reference GPU walk (trace.wgsl traverse) over
 * the GPU node layout, the CPU walks over the binned-SAH tree (engine/rt/bvh.c), and a brute
 * force over every triangle. TEST INFRASTRUCTURE ONLY (see lro.h).
 *
 * The triangle test is Moller-Trumbore exactly as trace.wgsl:159-173 and intersect.h:96-112
 * write it (identical sequences of f32 operations: cross and dot left to right, one division
 * for 1 / det, no fused multiply-add). */
#include "lro.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } V3;
static inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float fabs_(float x) { return x < 0 ? -x : x; }
static inline float wmin(float a, float b) { return b < a ? b : a; }
static inline float wmax(float a, float b) { return b > a ? b : a; }

/* hit_tri (trace.wgsl:159-173) = lr_hit_triangle_t (intersect.h:96-112): 1 and (t, b1, b2) on
 * a hit with t in (tmin, tmax) */
static int hit_tri(const float *tr, V3 o, V3 d, float tmin, float tmax, float *t_out, float *u_out, float *v_out)
{
    V3 q = v3(tr[0], tr[1], tr[2]), u = v3(tr[4], tr[5], tr[6]), v = v3(tr[8], tr[9], tr[10]);
    V3 pvec = vcross(d, v);
    float det = vdot(u, pvec);
    if (fabs_(det) < 1e-12f) return 0;
    float inv = 1.0f / det;
    V3 tvec = vsub(o, q);
    float b1 = vdot(tvec, pvec) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return 0;
    V3 qvec = vcross(tvec, u);
    float b2 = vdot(d, qvec) * inv;
    if (b2 < 0.0f || b1 + b2 > 1.0f) return 0;
    float t = vdot(v, qvec) * inv;
    if (t <= tmin || t >= tmax) return 0;
    *t_out = t; *u_out = b1; *v_out = b2;
    return 1;
}

/* safe_inv (media.wgsl:45-48) */
static inline float safe_inv1(float d) { return fabs_(d) > 1e-12f ? 1.0f / d : 1e30f; }

/* box_t (trace.wgsl:175-184): the entry distance, -1 on a miss */
static inline float box_t(const float *lo, const float *hi, V3 o, V3 inv, float tmin, float tmax)
{
    float t0x = (lo[0] - o.x) * inv.x, t0y = (lo[1] - o.y) * inv.y, t0z = (lo[2] - o.z) * inv.z;
    float t1x = (hi[0] - o.x) * inv.x, t1y = (hi[1] - o.y) * inv.y, t1z = (hi[2] - o.z) * inv.z;
    float ax = wmin(t0x, t1x), ay = wmin(t0y, t1y), az = wmin(t0z, t1z);
    float bx = wmax(t0x, t1x), by = wmax(t0y, t1y), bz = wmax(t0z, t1z);
    float enter = wmax(wmax(ax, ay), wmax(az, tmin));
    float leave = wmin(wmin(bx, by), wmin(bz, tmax));
    return leave < enter ? -1.0f : enter;
}

void lro_trace_gpu_layout(const uint32_t *nodes, uint32_t root, const uint32_t *aux, const float *tris,
                          const float *rays, int nrays, int any, LroHit *hits, LroCount *count)
{
    size_t cap = 256;
    uint32_t *stack = malloc(cap * 4);
    LroCount c = {0, 0};
    for (int r = 0; r < nrays; r++) {
        const float *ry = rays + 8 * (size_t)r;
        V3 o = v3(ry[0], ry[1], ry[2]), d = v3(ry[4], ry[5], ry[6]);
        float tmin = ry[3], best_t = ry[7];
        uint32_t best_id = 0xffffffffu;
        float bu = 0, bv = 0;
        V3 inv = v3(safe_inv1(d.x), safe_inv1(d.y), safe_inv1(d.z));
        size_t sp = 0;
        stack[sp++] = root;
        while (sp > 0) {
            uint32_t ni = stack[--sp];
            const uint32_t *nd = nodes + 8 * (size_t)ni;
            c.nodes++;
            if (nd[7] & 0x80000000u) {                              /* a leaf: its triangles */
                uint32_t cnt = nd[7] & 0x7fffffffu;
                int stop = 0;
                for (uint32_t k = 0; k < cnt; k++) {
                    uint32_t sid = aux[nd[3] + k];
                    float t, u, v;
                    c.tris++;
                    if (hit_tri(tris + 12 * (size_t)sid, o, d, tmin, best_t, &t, &u, &v)) {
                        best_t = t; best_id = sid; bu = u; bv = v;
                        if (any) { stop = 1; break; }
                    }
                }
                if (stop) break;
                continue;
            }
            /* inner: both children's boxes, the farther pushed first (trace.wgsl:374-389) */
            const float *na = (const float *)(nodes + 8 * (size_t)nd[3]);
            const float *nb = (const float *)(nodes + 8 * (size_t)nd[7]);
            c.nodes += 2;
            float tl = box_t(na, na + 4, o, inv, tmin, best_t);
            float tr = box_t(nb, nb + 4, o, inv, tmin, best_t);
            if (sp + 2 > cap) { cap *= 2; stack = realloc(stack, cap * 4); }
            if (tl >= 0.0f && tr >= 0.0f) {
                int left_first = tl <= tr;
                stack[sp++] = left_first ? nd[7] : nd[3];
                stack[sp++] = left_first ? nd[3] : nd[7];
            } else if (tl >= 0.0f) stack[sp++] = nd[3];
            else if (tr >= 0.0f) stack[sp++] = nd[7];
        }
        hits[r].t = best_t; hits[r].id = best_id; hits[r].u = bu; hits[r].v = bv;
    }
    free(stack);
    if (count) { count->nodes += c.nodes; count->tris += c.tris; }
}

/* ---- the CPU walks (bvh.c:541-663) ---- */
/* aabb_hit (intersect.h:38-51) */
static int aabb_hit(const float *lo, const float *hi, V3 o, V3 inv, float tmin, float tmax, float *te)
{
    const float oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
    for (int k = 0; k < 3; k++) {
        float t0 = (lo[k] - oo[k]) * ii[k], t1 = (hi[k] - oo[k]) * ii[k];
        if (ii[k] < 0) { float t = t0; t0 = t1; t1 = t; }
        tmin = t0 != t0 ? tmin : (t0 > tmin ? t0 : tmin);         /* r_fmax(t0, tmin) */
        tmax = tmax != tmax ? t1 : (t1 < tmax ? t1 : tmax);        /* r_fmin(t1, tmax) */
    }
    if (tmax < tmin) return 0;
    *te = tmin;
    return 1;
}

static void sah_one(const LroSah *b, const float *tris, const float *ry, int any, LroHit *h)
{
    V3 o = v3(ry[0], ry[1], ry[2]), d = v3(ry[4], ry[5], ry[6]);
    float tmin = ry[3], best = ry[7];
    int hit_id = -1;
    float ku = 0, kv = 0;
    h->t = best; h->id = 0xffffffffu; h->u = h->v = 0;
    if (b->nnodes == 0) return;
    if (b->wide) {
        /* wide_intersect (bvh.c:543-604) */
        const float ox = o.x, oy = o.y, oz = o.z;
        const float ix = 1.0f / d.x, iy = 1.0f / d.y, iz = 1.0f / d.z;
        const int nx = ix < 0 ? 3 : 0, ny = iy < 0 ? 4 : 1, nz = iz < 0 ? 5 : 2;
        const int fx = 3 - nx, fy = 5 - ny, fz = 7 - nz;
        struct { int node, count; float t; } stack[4 * 100 + 8];
        int sp = 0;
        stack[sp].node = 0; stack[sp].count = 0; stack[sp].t = tmin; sp++;
        while (sp > 0) {
            sp--;
            if (stack[sp].t > best) continue;
            if (stack[sp].count > 0) {
                const int start = stack[sp].node, cnt = stack[sp].count;
                for (int k = 0; k < cnt; k++) {
                    int id = b->refs[start + k];
                    float t, u, v;
                    if (hit_tri(tris + 12 * (size_t)id, o, d, tmin, best, &t, &u, &v)) {
                        best = t; hit_id = id; ku = u; kv = v;
                        if (any) goto done;
                    }
                }
                continue;
            }
            const LroWideNode *n = &b->wide[stack[sp].node];
            float tn[4];
            int hit[4];
            for (int k = 0; k < 4; k++) {
                float a = (n->b[nx][k] - ox) * ix, c = (n->b[fx][k] - ox) * ix;
                float lo = a > tmin ? a : tmin, hi = c < best ? c : best;
                a = (n->b[ny][k] - oy) * iy; c = (n->b[fy][k] - oy) * iy;
                lo = a > lo ? a : lo; hi = c < hi ? c : hi;
                a = (n->b[nz][k] - oz) * iz; c = (n->b[fz][k] - oz) * iz;
                lo = a > lo ? a : lo; hi = c < hi ? c : hi;
                tn[k] = lo;
                hit[k] = lo <= hi;
            }
            int order[4], m = 0;
            for (int k = 0; k < 4; k++) {
                if (!hit[k] || n->count[k] < 0) continue;
                int j = m++;
                while (j > 0 && tn[order[j - 1]] < tn[k]) { order[j] = order[j - 1]; j--; }
                order[j] = k;
            }
            for (int q = 0; q < m; q++) {
                int k = order[q];
                stack[sp].node = n->child[k]; stack[sp].count = n->count[k]; stack[sp].t = tn[k];
                sp++;
            }
        }
    } else {
        /* bvh_intersect, binary (bvh.c:608-663) */
        V3 inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        struct { int node; float t; } stack[2 * 100 + 8];
        int sp = 0;
        float te;
        if (!aabb_hit(b->nodes[0].lo, b->nodes[0].hi, o, inv, tmin, best, &te)) goto done;
        stack[sp].node = 0; stack[sp].t = te; sp++;
        while (sp > 0) {
            sp--;
            if (stack[sp].t > best) continue;
            const LroSahNode *n = &b->nodes[stack[sp].node];
            if (n->count > 0) {
                for (int k = 0; k < n->count; k++) {
                    int id = b->refs[n->start + k];
                    float t, u, v;
                    if (hit_tri(tris + 12 * (size_t)id, o, d, tmin, best, &t, &u, &v)) {
                        best = t; hit_id = id; ku = u; kv = v;
                        if (any) goto done;
                    }
                }
                continue;
            }
            float tl = 0, tr = 0;
            int hl = aabb_hit(b->nodes[n->left].lo, b->nodes[n->left].hi, o, inv, tmin, best, &tl);
            int hr = aabb_hit(b->nodes[n->right].lo, b->nodes[n->right].hi, o, inv, tmin, best, &tr);
            if (hl && hr) {
                int near = tl <= tr ? n->left : n->right, far = tl <= tr ? n->right : n->left;
                float tnear = tl < tr ? tl : tr, tfar = tl > tr ? tl : tr;
                stack[sp].node = far; stack[sp].t = tfar; sp++;
                stack[sp].node = near; stack[sp].t = tnear; sp++;
            } else if (hl) { stack[sp].node = n->left; stack[sp].t = tl; sp++; }
            else if (hr) { stack[sp].node = n->right; stack[sp].t = tr; sp++; }
        }
    }
done:
    if (hit_id >= 0) { h->t = best; h->id = (uint32_t)hit_id; h->u = ku; h->v = kv; }
}

void lro_trace_sah(const LroSah *b, const float *tris, const float *rays, int nrays, int any, LroHit *hits)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int r = 0; r < nrays; r++) sah_one(b, tris, rays + 8 * (size_t)r, any, &hits[r]);
}

void lro_trace_brute(const float *tris, int ntris, const float *rays, int nrays, int any, LroHit *hits, float *second)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int r = 0; r < nrays; r++) {
        const float *ry = rays + 8 * (size_t)r;
        V3 o = v3(ry[0], ry[1], ry[2]), d = v3(ry[4], ry[5], ry[6]);
        float tmin = ry[3], best = ry[7], sec = ry[7];
        uint32_t id = 0xffffffffu;
        float bu = 0, bv = 0;
        for (int i = 0; i < ntris; i++) {
            float t, u, v;
            /* every candidate in (tmin, tmax), to know the second nearest too */
            if (!hit_tri(tris + 12 * (size_t)i, o, d, tmin, ry[7], &t, &u, &v)) continue;
            if (t < best) {
                sec = best;
                best = t; id = (uint32_t)i; bu = u; bv = v;
                if (any) break;
            } else if (t < sec) sec = t;
        }
        hits[r].t = best; hits[r].id = id; hits[r].u = bu; hits[r].v = bv;
        if (second) second[r] = sec;
    }
}
