/* lro_build.c - the oracle's BVH builders: the reference GPU's LBVH and PLOC rules and the
 * CPU's binned-SAH tree, restated in plain C. TEST INFRASTRUCTURE ONLY (see lro.h). */
#include "lro.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

/* ---- PCG32 (engine/base/rng.h:13-37) ---- */
uint32_t lro_rng_u32(LroRng *r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

void lro_rng_seed(LroRng *r, uint64_t seed, uint64_t stream)
{
    r->state = 0;
    r->inc = (stream << 1u) | 1u;
    lro_rng_u32(r);
    r->state += seed;
    lro_rng_u32(r);
}

float lro_rng_float(LroRng *r) { return (float)(lro_rng_u32(r) >> 8) * (1.0f / 16777216.0f); }
double lro_rng_double(LroRng *r) { return lro_rng_u32(r) * (1.0 / 4294967296.0); }

/* WGSL's min / max on f32 (the builders' boxes): e2 if e2 < e1 (resp. >), else e1 */
static inline float wmin(float a, float b) { return b < a ? b : a; }
static inline float wmax(float a, float b) { return b > a ? b : a; }
/* r_fmin / r_fmax (lrmath.h:56-57): the other number when b is a NaN */
static inline float rfmin(float a, float b) { return b != b ? a : (a < b ? a : b); }
static inline float rfmax(float a, float b) { return b != b ? a : (a > b ? a : b); }

/* ---- boxes (scene.c:41-74) ---- */
void lro_tri_boxes(const float *tris, int n, float *boxes)
{
    for (int i = 0; i < n; i++) {
        const float *t = tris + 12 * (size_t)i;
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            float q = t[k], a = q + t[4 + k], b = q + t[8 + k];   /* q, q + u, q + v */
            lo[k] = q; hi[k] = q;
            lo[k] = rfmin(lo[k], a); hi[k] = rfmax(hi[k], a);      /* grow_box (scene.c:41-44) */
            lo[k] = rfmin(lo[k], b); hi[k] = rfmax(hi[k], b);
        }
        float *o = boxes + 8 * (size_t)i;
        for (int k = 0; k < 3; k++) {
            o[k] = lo[k] - 1e-3f;                                  /* the pad (scene.c:71-73) */
            o[4 + k] = hi[k] + 1e-3f;
        }
        o[3] = 0.0f; o[7] = 0.0f;
    }
}

void lro_union(const float *boxes, int n, float lo[3], float hi[3])
{
    for (int k = 0; k < 3; k++) { lo[k] = 1e30f; hi[k] = -1e30f; }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = rfmin(lo[k], boxes[8 * (size_t)i + k]);
            hi[k] = rfmax(hi[k], boxes[8 * (size_t)i + 4 + k]);
        }
}

void lro_inv_extent(const float lo[3], const float hi[3], float ext[3])
{
    for (int k = 0; k < 3; k++) {
        float e = hi[k] - lo[k];                                   /* v_sub in real = float */
        ext[k] = e > 0 ? (float)(1.0 / (double)e) : 0.0f;          /* trace_gpu.c:609-611 */
    }
}

/* ---- Morton codes (trace_build.wgsl:16-33) ---- */
static uint32_t spread(uint32_t v)
{
    uint32_t x = v & 1023u;
    x = (x | (x << 16u)) & 0x030000ffu;
    x = (x | (x << 8u)) & 0x0300f00fu;
    x = (x | (x << 4u)) & 0x030c30c3u;
    x = (x | (x << 2u)) & 0x09249249u;
    return x;
}

static float clamp01(float x) { return wmin(wmax(x, 0.0f), 0.999f); }   /* WGSL clamp(x, 0, .999) */

void lro_morton(const float *boxes, int n, const float lo[3], const float ext[3], uint32_t *codes)
{
    for (int i = 0; i < n; i++) {
        const float *b = boxes + 8 * (size_t)i;
        uint32_t q[3];
        for (int k = 0; k < 3; k++) {
            float c = (b[k] + b[4 + k]) * 0.5f;
            q[k] = (uint32_t)(clamp01((c - lo[k]) * ext[k]) * 1024.0f);
        }
        codes[i] = (spread(q[0]) << 2u) | (spread(q[1]) << 1u) | spread(q[2]);
    }
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* bsort (trace_build.wgsl:35-50) sorts the keys (code, item) ascending; the padding keys
 * (0xffffffff, 0xffffffff) sort last, so the first n are the items' keys in order */
void lro_sort(const uint32_t *codes, int n, uint32_t *sorted_codes, uint32_t *sorted_items)
{
    uint64_t *k = malloc((size_t)(n > 0 ? n : 1) * 8);
    for (int i = 0; i < n; i++) k[i] = ((uint64_t)codes[i] << 32) | (uint32_t)i;
    qsort(k, (size_t)n, 8, cmp_u64);
    for (int i = 0; i < n; i++) { sorted_codes[i] = (uint32_t)(k[i] >> 32); sorted_items[i] = (uint32_t)k[i]; }
    free(k);
}

/* ---- LBVH (trace_build.wgsl:52-122) ---- */
static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* delta (trace_build.wgsl:52-58): common prefix of keys i and j; equal codes: 32 + that of
 * their positions; j out of range: -1 */
static int delta(const uint32_t *codes, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

static void put_node(uint32_t *nd, const float *lo, uint32_t a, const float *hi, uint32_t b)
{
    memcpy(nd, lo, 12); nd[3] = a;
    memcpy(nd + 4, hi, 12); nd[7] = b;
}

/* leaves (trace_build.wgsl:98-107) */
static void lbvh_leaves(const uint32_t *items, const float *boxes, int n, uint32_t ref0, uint32_t ref_add,
                        uint32_t *nodes, uint32_t *aux)
{
    for (int j = 0; j < n; j++) {
        uint32_t item = items[j];
        const float *b = boxes + 8 * (size_t)item;
        put_node(nodes + 8 * (size_t)(n - 1 + j), b, ref0 + (uint32_t)j, b + 4, 0x80000001u);
        aux[j] = item + ref_add;
    }
}

/* a node's box from its children's (up, trace_build.wgsl:109-122): min / max of left, right */
static void join_boxes(uint32_t *nd, const uint32_t *l, const uint32_t *r)
{
    float a[8], b[8], lo[3], hi[3];
    memcpy(a, l, 32); memcpy(b, r, 32);
    for (int k = 0; k < 3; k++) { lo[k] = wmin(a[k], b[k]); hi[k] = wmax(a[4 + k], b[4 + k]); }
    memcpy(nd, lo, 12); memcpy(nd + 4, hi, 12);
}

/* boxes of the internal nodes bottom-up, and the depth (root = 1, as check_tree counts) */
static int refit(uint32_t *nodes, uint32_t node0, uint32_t i)
{
    uint32_t *nd = nodes + 8 * (size_t)i;
    if (nd[7] & LRO_LEAF) return 1;
    uint32_t a = nd[3] - node0, b = nd[7] - node0;
    int da = refit(nodes, node0, a), db = refit(nodes, node0, b);
    join_boxes(nd, nodes + 8 * (size_t)a, nodes + 8 * (size_t)b);
    return 1 + (da > db ? da : db);
}

int lro_lbvh(const uint32_t *codes, const uint32_t *items, const float *boxes, int n,
             uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux)
{
    if (n <= 0) return 0;
    /* hier (trace_build.wgsl:60-96), one internal node i at a time */
    for (int i = 0; i < n - 1; i++) {
        int d = delta(codes, n, i, i + 1) > delta(codes, n, i, i - 1) ? 1 : -1;
        int dmin = delta(codes, n, i, i - d);
        long lmax = 2;
        for (int g = 0; g < 32; g++) {
            if (lmax >= 2L * n || delta(codes, n, i, (int)(i + lmax * d)) <= dmin) break;
            lmax *= 2;
        }
        long l = 0, t = lmax / 2;
        for (int g = 0; g < 32; g++) {
            if (t < 1) break;
            if (delta(codes, n, i, (int)(i + (l + t) * d)) > dmin) l += t;
            t /= 2;
        }
        int j = (int)(i + l * d);
        int dnode = delta(codes, n, i, j);
        long s = 0, div = 2;
        for (int g = 0; g < 32; g++) {
            long st = (l + div - 1) / div;
            if (delta(codes, n, i, (int)(i + (s + st) * d)) > dnode) s += st;
            if (st <= 1) break;
            div *= 2;
        }
        int gam = (int)(i + s * d + (d < 0 ? d : 0));
        int mn = i < j ? i : j, mx = i > j ? i : j;
        int left = mn == gam ? n - 1 + gam : gam;
        int right = mx == gam + 1 ? n + gam : gam + 1;
        nodes[8 * (size_t)i + 3] = node0 + (uint32_t)left;
        nodes[8 * (size_t)i + 7] = node0 + (uint32_t)right;
    }
    lbvh_leaves(items, boxes, n, ref0, ref_add, nodes, aux);
    return refit(nodes, node0, 0);
}

/* The binary radix tree again, bottom-up (Apetrei 2014, "Fast and Simple Agglomerative LBVH
 * Construction"): a node over keys [l, r] has its parent's split at r when the keys r, r+1
 * share the longer prefix (else at l - 1). Karras's numbering follows from it: an internal
 * left child is numbered by its last key, an internal right child by its first, the root 0.
 * This is the construction the device uses (one pass, no parent pointers). */
static int karras_index(const uint32_t *codes, int n, int l, int r)
{
    if (l == 0 && r == n - 1) return 0;
    int left_child = l == 0 || (r != n - 1 && delta(codes, n, r, r + 1) > delta(codes, n, l - 1, l));
    return left_child ? r : l;
}

int lro_lbvh_bottom_up(const uint32_t *codes, const uint32_t *items, const float *boxes, int n,
                       uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux)
{
    if (n <= 0) return 0;
    lbvh_leaves(items, boxes, n, ref0, ref_add, nodes, aux);
    if (n == 1) return 1;
    int *other = malloc((size_t)(n - 1) * sizeof(int));
    int *height = calloc((size_t)(2 * n - 1), sizeof(int));
    for (int k = 0; k < n - 1; k++) other[k] = -1;
    for (int j = 0; j < n; j++) height[n - 1 + j] = 1;
    int depth = 1;
    for (int j = 0; j < n; j++) {
        int l = j, r = j;
        for (;;) {
            if (l == 0 && r == n - 1) break;
            int left_child = l == 0 || (r != n - 1 && delta(codes, n, r, r + 1) > delta(codes, n, l - 1, l));
            int p = left_child ? r : l - 1;              /* the parent's split */
            int end = left_child ? l : r;                /* this side's far end */
            if (other[p] < 0) { other[p] = end; break; } /* first to arrive: the other side goes on */
            int L = left_child ? end : other[p], R = left_child ? other[p] : end;
            int lc = L == p ? n - 1 + p : p, rc = p + 1 == R ? n + p : p + 1;
            int me = karras_index(codes, n, L, R);
            uint32_t *nd = nodes + 8 * (size_t)me;
            nd[3] = node0 + (uint32_t)lc;
            nd[7] = node0 + (uint32_t)rc;
            join_boxes(nd, nodes + 8 * (size_t)lc, nodes + 8 * (size_t)rc);
            height[me] = 1 + (height[lc] > height[rc] ? height[lc] : height[rc]);
            if (height[me] > depth) depth = height[me];
            l = L; r = R;
        }
    }
    free(other);
    free(height);
    return depth;
}

/* ---- PLOC (trace_ploc.wgsl:1-301), radius 16 ---- */
#define PLOC_R 16

/* area (trace_ploc.wgsl:18): half the surface, evaluated as written, left to right, with no
 * fused multiply-add (the one evaluation order the device also uses) */
static float half_area(const float *lo, const float *hi)
{
    float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    float a = ex * ey, b = ey * ez, c = ez * ex;
    float s = a + b;
    return s + c;
}

/* nearer (trace_ploc.wgsl:60-66): a smaller distance, or on a tie the lower pair of positions */
static int nearer(float d, uint32_t i, uint32_t j, float bd, uint32_t bj)
{
    if (bj == 0xffffffffu || d < bd) return 1;
    if (d > bd) return 0;
    uint32_t mn = i < j ? i : j, bmn = i < bj ? i : bj;
    uint32_t mx = i > j ? i : j, bmx = i > bj ? i : bj;
    return mn < bmn || (mn == bmn && mx < bmx);
}

int lro_ploc(const uint32_t *codes, const uint32_t *items, const float *boxes, int n,
             uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux, int *depth,
             uint64_t *live)
{
    if (live) *live = 0;
    if (n <= 2) {                                 /* build_queue: PLOC only above 2 (trace_gpu.c:615) */
        int d = lro_lbvh(codes, items, boxes, n, node0, ref0, ref_add, nodes, aux);
        if (depth) *depth = d;
        return 0;
    }
    /* leaves (trace_build.wgsl:98-107), then init: the clusters are the leaves n-1+j */
    lbvh_leaves(items, boxes, n, ref0, ref_add, nodes, aux);
    size_t nx = 2 * (size_t)n - 1;
    float *bx = malloc(nx * 6 * sizeof(float));   /* the clusters' boxes (pl's last part) */
    uint32_t *list = malloc((size_t)n * 4), *next = malloc((size_t)n * 4), *nn = malloc((size_t)n * 4);
    int *height = calloc(nx, sizeof(int));
    for (int j = 0; j < n; j++) {
        const float *f = (const float *)(nodes + 8 * (size_t)(n - 1 + j));
        memcpy(bx + 6 * (size_t)(n - 1 + j), f, 12);
        memcpy(bx + 6 * (size_t)(n - 1 + j) + 3, f + 4, 12);
        list[j] = (uint32_t)(n - 1 + j);
        height[n - 1 + j] = 1;
    }
    uint32_t nc = (uint32_t)n;
    int rounds = 0;
    while (nc > 1) {
        if (live) *live += nc;
        /* nn (trace_ploc.wgsl:99-128) */
        for (uint32_t i = 0; i < nc; i++) {
            const float *bi = bx + 6 * (size_t)list[i];
            uint32_t j0 = i > PLOC_R ? i - PLOC_R : 0, j1 = i + PLOC_R < nc - 1 ? i + PLOC_R : nc - 1;
            float bd = 0.0f;
            uint32_t bj = 0xffffffffu;
            for (uint32_t j = j0; j <= j1; j++) {
                if (j == i) continue;
                const float *bjx = bx + 6 * (size_t)list[j];
                float lo[3], hi[3];
                for (int k = 0; k < 3; k++) { lo[k] = wmin(bi[k], bjx[k]); hi[k] = wmax(bi[3 + k], bjx[3 + k]); }
                float d = half_area(lo, hi);
                if (nearer(d, i, j, bd, bj)) { bd = d; bj = j; }
            }
            nn[i] = bj;
        }
        /* keep_merge (trace_ploc.wgsl:140-145), the scans and scatter (:214-234), in order */
        uint32_t out = 0, merges = 0;
        for (uint32_t i = 0; i < nc; i++) {
            uint32_t j = nn[i];
            int mu = j < nc && nn[j] == i;
            int keep = !(mu && i > j), merge = mu && i < j;
            if (!keep) continue;
            uint32_t c = list[i], o = c;
            if (merge) {
                uint32_t d = list[j];
                uint32_t k = ((uint32_t)n - 2u) - ((uint32_t)n - nc + merges);
                merges++;
                /* merge (trace_ploc.wgsl:39-46) */
                float *bk = bx + 6 * (size_t)k;
                const float *bc = bx + 6 * (size_t)c, *bdd = bx + 6 * (size_t)d;
                for (int q = 0; q < 3; q++) { bk[q] = wmin(bc[q], bdd[q]); bk[3 + q] = wmax(bc[3 + q], bdd[3 + q]); }
                uint32_t xs[2] = {c, d};
                for (int s = 0; s < 2; s++) {
                    uint32_t x = xs[s];
                    uint32_t *slot = nodes + 8 * (size_t)(1 + 2 * k + (uint32_t)s);
                    const float *b = bx + 6 * (size_t)x;
                    /* final_node (trace_ploc.wgsl:33-38) */
                    if (x < (uint32_t)n - 1u) put_node(slot, b, node0 + 1u + 2u * x, b + 3, node0 + 2u + 2u * x);
                    else put_node(slot, b, ref0 + x - ((uint32_t)n - 1u), b + 3, 0x80000001u);
                }
                height[k] = 1 + (height[c] > height[d] ? height[c] : height[d]);
                if (k == 0) put_node(nodes, bk, node0 + 1u, bk + 3, node0 + 2u);
                o = k;
            }
            next[out++] = o;
        }
        uint32_t *t = list; list = next; next = t;
        nc = out;
        rounds++;
    }
    if (depth) *depth = height[0];
    free(bx); free(list); free(next); free(nn); free(height);
    return rounds;
}

/* ---- check_tree (trace_gpu.c:766-805) ---- */
int lro_check_tree(const uint32_t *w, int n, uint32_t node0, double *cost, int *depth)
{
    if (n <= 0) return 0;
    size_t nn = 2 * (size_t)n - 1;
    unsigned char *seen = calloc(nn, 1);
    uint32_t *stack = malloc(nn * 2 * sizeof(uint32_t));
    float f[8];
    double sum = 0, root = 0;
    int leaves = 0, ok = 1, top = 0, deep = 0;
    stack[top++] = 0; stack[top++] = 1;
    seen[0] = 1;
    while (top > 0 && ok) {
        uint32_t d = stack[--top], i = stack[--top];
        if ((int)d > deep) deep = (int)d;
        memcpy(f, w + 8 * (size_t)i, 32);
        double ex = (double)f[4] - f[0], ey = (double)f[5] - f[1], ez = (double)f[6] - f[2];
        double area = 2 * (ex * ey + ey * ez + ez * ex);
        if (!(area >= 0)) ok = 0;
        sum += area;
        if (i == 0) root = area;
        uint32_t a = w[8 * (size_t)i + 3], b = w[8 * (size_t)i + 7];
        if (b & 0x80000000u) { leaves++; continue; }
        uint32_t c[2] = {a - node0, b - node0};
        for (int k = 0; k < 2; k++) {
            if (c[k] >= nn || seen[c[k]]) { ok = 0; break; }
            seen[c[k]] = 1;
            float g[8];
            memcpy(g, w + 8 * (size_t)c[k], 32);
            for (int x = 0; x < 3; x++)
                if (g[x] < f[x] - 1e-3f * (1 + __builtin_fabsf(f[x])) || g[4 + x] > f[4 + x] + 1e-3f * (1 + __builtin_fabsf(f[4 + x]))) ok = 0;
            stack[top++] = c[k]; stack[top++] = d + 1;
        }
    }
    free(seen);
    free(stack);
    if (cost) *cost = root > 0 ? sum / root : 0;
    if (depth) *depth = deep;
    return ok && leaves == n;
}

int lro_build_tris(const float *tris, int n, int ploc, uint32_t node0, uint32_t ref0, uint32_t ref_add,
                   uint32_t *nodes, uint32_t *aux, uint32_t *keys_out)
{
    if (n <= 0) return 0;
    float *boxes = malloc((size_t)n * 32);
    uint32_t *codes = malloc((size_t)n * 4), *sc = malloc((size_t)n * 4), *si = malloc((size_t)n * 4);
    float lo[3], hi[3], ext[3];
    lro_tri_boxes(tris, n, boxes);
    lro_union(boxes, n, lo, hi);
    lro_inv_extent(lo, hi, ext);
    lro_morton(boxes, n, lo, ext, codes);
    lro_sort(codes, n, sc, si);
    int depth = 0;
    if (ploc) lro_ploc(sc, si, boxes, n, node0, ref0, ref_add, nodes, aux, &depth, NULL);
    else depth = lro_lbvh(sc, si, boxes, n, node0, ref0, ref_add, nodes, aux);
    if (keys_out) { memcpy(keys_out, sc, (size_t)n * 4); memcpy(keys_out + n, si, (size_t)n * 4); }
    free(boxes); free(codes); free(sc); free(si);
    return depth;
}

/* ---- the CPU's binned-SAH BVH (engine/rt/bvh.c), float build ---- */
typedef struct { float lo[3], hi[3]; } Box;
#define SAH_MAX_DEPTH 100          /* BVH_MAX_DEPTH (bvh.h:17) */
#define SAH_MAX_BINS 256

static Box box_empty(void)
{
    Box b = {{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    return b;
}
static inline float mn(float a, float b) { return a < b ? a : b; }   /* bvh.c:17-18 */
static inline float mx(float a, float b) { return a > b ? a : b; }
static void box_add_box(Box *b, const Box *o)
{
    for (int k = 0; k < 3; k++) { b->lo[k] = mn(b->lo[k], o->lo[k]); b->hi[k] = mx(b->hi[k], o->hi[k]); }
}
static void box_add_point(Box *b, const float *p)
{
    for (int k = 0; k < 3; k++) { b->lo[k] = mn(b->lo[k], p[k]); b->hi[k] = mx(b->hi[k], p[k]); }
}
static float box_area(const Box *b)                                  /* bvh.c:30-34 */
{
    float dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0;
    float s = dx * dy + dy * dz;
    return 2.0f * (s + dz * dx);
}

typedef struct {
    const Box *box;
    float (*cen)[3];
    int *refs;
    int max_leaf, bins;
    LroSahNode *nodes; int nnodes;
    int leaves, depth;
} SahBuild;

static void range_bounds(const SahBuild *sb, int start, int count, Box *bo, Box *co)
{
    Box b = box_empty(), c = box_empty();
    for (int i = start; i < start + count; i++) {
        box_add_box(&b, &sb->box[sb->refs[i]]);
        box_add_point(&c, sb->cen[sb->refs[i]]);
    }
    *bo = b; *co = c;
}

/* sah_split (bvh.c:146-223) */
static float sah_split(const SahBuild *sb, int start, int count, const Box *node_box, const Box *cbox,
                       int *best_axis, float *best_cost)
{
    int nb = sb->bins < 2 ? 2 : (sb->bins > SAH_MAX_BINS ? SAH_MAX_BINS : sb->bins);
    float parent_area = box_area(node_box), best_pos = 0;
    *best_cost = FLT_MAX;
    *best_axis = -1;
    int use[3];
    float los[3], scale[3];
    for (int a = 0; a < 3; a++) {
        float lo = cbox->lo[a], extent = cbox->hi[a] - lo;
        use[a] = extent > 1e-12f;
        los[a] = lo;
        scale[a] = use[a] ? (float)nb / extent : 0;
    }
    static _Thread_local int bn[3][SAH_MAX_BINS];
    static _Thread_local Box bbox[3][SAH_MAX_BINS];
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < nb; k++) { bn[a][k] = 0; bbox[a][k] = box_empty(); }
    for (int i = start; i < start + count; i++) {                  /* range_bins (bvh.c:112-128) */
        int p = sb->refs[i];
        for (int a = 0; a < 3; a++) {
            if (!use[a]) continue;
            float f = (sb->cen[p][a] - los[a]) * scale[a];
            int k = f > 0 ? (f < nb ? (int)f : nb - 1) : 0;
            bn[a][k]++;
            box_add_box(&bbox[a][k], &sb->box[p]);
        }
    }
    for (int a = 0; a < 3; a++) {
        if (!use[a]) continue;
        const float lo = los[a], extent = cbox->hi[a] - lo;
        float right_area[SAH_MAX_BINS];
        int right_n[SAH_MAX_BINS];
        Box acc = box_empty();
        int n = 0;
        for (int k = nb - 1; k > 0; k--) {
            box_add_box(&acc, &bbox[a][k]);
            n += bn[a][k];
            right_area[k] = box_area(&acc);
            right_n[k] = n;
        }
        acc = box_empty();
        n = 0;
        for (int k = 1; k < nb; k++) {
            box_add_box(&acc, &bbox[a][k - 1]);
            n += bn[a][k - 1];
            if (n == 0 || right_n[k] == 0) continue;
            float l = box_area(&acc) * (float)n, r = right_area[k] * (float)right_n[k];
            float cost = 1.0f + (l + r) / parent_area;
            if (cost < *best_cost) {
                *best_cost = cost;
                *best_axis = a;
                best_pos = lo + extent * (float)k / (float)nb;
            }
        }
    }
    return best_pos;
}

/* partition_bounds (bvh.c:73-91) */
static int partition_bounds(SahBuild *sb, int start, int count, int axis, float pos, Box kids[4])
{
    Box lb = box_empty(), lc = box_empty(), rb = box_empty(), rc = box_empty();
    int i = start, j = start + count - 1;
    while (i <= j) {
        int p = sb->refs[i];
        if (sb->cen[p][axis] < pos) { box_add_box(&lb, &sb->box[p]); box_add_point(&lc, sb->cen[p]); i++; }
        else { sb->refs[i] = sb->refs[j]; sb->refs[j] = p; j--; box_add_box(&rb, &sb->box[p]); box_add_point(&rc, sb->cen[p]); }
    }
    kids[0] = lb; kids[1] = lc; kids[2] = rb; kids[3] = rc;
    return i - start;
}

/* choose_split (bvh.c:230-283), SAH, serial (the same decisions as the parallel build) */
static int choose_split(SahBuild *sb, int start, int count, int depth, Box *box_out, const Box *have, Box *kids)
{
    Box box, cbox;
    if (have) { box = have[0]; cbox = have[1]; }
    else range_bounds(sb, start, count, &box, &cbox);
    *box_out = box;
    int nleft = 0, kids_done = 0;
    if (count > 1 && depth < SAH_MAX_DEPTH) {
        int axis = -1;
        float cost;
        float pos = sah_split(sb, start, count, &box, &cbox, &axis, &cost);
        if (!(axis >= 0 && (cost < (float)count || count > sb->max_leaf))) axis = -1;
        if (axis >= 0) { nleft = partition_bounds(sb, start, count, axis, pos, kids); kids_done = 1; }
        if ((nleft == 0 || nleft == count) && count > sb->max_leaf) { nleft = count / 2; kids_done = 0; }
        if (nleft == count) nleft = 0;
    }
    if (nleft > 0 && !kids_done) {
        range_bounds(sb, start, nleft, &kids[0], &kids[1]);
        range_bounds(sb, start + nleft, count - nleft, &kids[2], &kids[3]);
    }
    return nleft;
}

/* build (bvh.c:288-306): preorder */
static int sah_node(SahBuild *sb, int start, int count, int depth, const Box *have)
{
    int node = sb->nnodes++;
    memset(&sb->nodes[node], 0, sizeof(LroSahNode));
    if (depth > sb->depth) sb->depth = depth;
    Box box, kids[4];
    int nleft = choose_split(sb, start, count, depth, &box, have, kids);
    memcpy(sb->nodes[node].lo, box.lo, 12);
    memcpy(sb->nodes[node].hi, box.hi, 12);
    if (nleft == 0) {
        sb->nodes[node].start = start;
        sb->nodes[node].count = count;
        sb->leaves++;
        return node;
    }
    int l = sah_node(sb, start, nleft, depth + 1, kids);
    int r = sah_node(sb, start + nleft, count - nleft, depth + 1, kids + 2);
    sb->nodes[node].left = l;
    sb->nodes[node].right = r;
    sb->nodes[node].count = 0;
    return node;
}

static float node_area(const LroSahNode *n)
{
    Box b;
    memcpy(b.lo, n->lo, 12); memcpy(b.hi, n->hi, 12);
    return box_area(&b);
}

/* collapse (bvh.c:497-532) */
static int collapse(LroSah *b, int node, int *cap)
{
    if (b->nwide >= *cap) { *cap = *cap ? *cap * 2 : 256; b->wide = realloc(b->wide, (size_t)*cap * sizeof(LroWideNode)); }
    int w = b->nwide++;
    int kids[4], n = 0;
    const LroSahNode *bn = &b->nodes[node];
    if (bn->count > 0) kids[n++] = node;
    else { kids[n++] = bn->left; kids[n++] = bn->right; }
    while (n < 4) {
        int best = -1;
        float area = -1;
        for (int k = 0; k < n; k++) {
            const LroSahNode *c = &b->nodes[kids[k]];
            float a = node_area(c);
            if (c->count == 0 && a > area) { area = a; best = k; }
        }
        if (best < 0) break;
        const LroSahNode *c = &b->nodes[kids[best]];
        kids[best] = c->left;
        kids[n++] = c->right;
    }
    int child[4], count[4];
    for (int k = 0; k < 4; k++) {
        if (k >= n) { child[k] = 0; count[k] = -1; continue; }
        const LroSahNode *c = &b->nodes[kids[k]];
        if (c->count > 0) { child[k] = c->start; count[k] = c->count; }
        else { child[k] = collapse(b, kids[k], cap); count[k] = 0; }
    }
    LroWideNode *o = &b->wide[w];
    for (int k = 0; k < 4; k++) {
        Box box = box_empty();
        if (k < n) { memcpy(box.lo, b->nodes[kids[k]].lo, 12); memcpy(box.hi, b->nodes[kids[k]].hi, 12); }
        for (int q = 0; q < 3; q++) { o->b[q][k] = box.lo[q]; o->b[3 + q][k] = box.hi[q]; }
        o->child[k] = child[k];
        o->count[k] = count[k];
    }
    return w;
}

LroSah *lro_sah_build(const float *boxes, int n, int max_leaf, int bins, int wide)
{
    LroSah *b = calloc(1, sizeof(LroSah));
    if (n <= 0) return b;
    SahBuild sb;
    memset(&sb, 0, sizeof(sb));
    Box *bx = malloc((size_t)n * sizeof(Box));
    float (*cen)[3] = malloc((size_t)n * sizeof(*cen));
    b->refs = malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; i++) {
        memcpy(bx[i].lo, boxes + 8 * (size_t)i, 12);
        memcpy(bx[i].hi, boxes + 8 * (size_t)i + 4, 12);
        for (int k = 0; k < 3; k++) cen[i][k] = (bx[i].lo[k] + bx[i].hi[k]) * 0.5f;   /* bvh.c:421 */
        b->refs[i] = i;
    }
    b->nrefs = n;
    sb.box = bx; sb.cen = cen; sb.refs = b->refs;
    sb.max_leaf = max_leaf < 1 ? 1 : max_leaf;
    sb.bins = bins;
    sb.nodes = malloc((2 * (size_t)n - 1) * sizeof(LroSahNode));
    sah_node(&sb, 0, n, 0, NULL);
    b->nodes = sb.nodes; b->nnodes = sb.nnodes;
    b->leaves = sb.leaves; b->depth = sb.depth;
    free(bx); free(cen);
    if (wide) { int cap = 0; collapse(b, 0, &cap); }
    return b;
}

void lro_sah_free(LroSah *b)
{
    if (!b) return;
    free(b->nodes); free(b->refs); free(b->wide); free(b);
}

void lro_sah_info(const LroSah *b, int *nnodes, int *nrefs, int *depth, int *nwide)
{
    *nnodes = b->nnodes; *nrefs = b->nrefs; *depth = b->depth; *nwide = b->nwide;
}

double lro_sah_cost(const LroSah *b)
{
    if (b->nnodes == 0) return 0;
    float root = node_area(&b->nodes[0]), sum = 0;
    for (int i = 0; i < b->nnodes; i++) sum += node_area(&b->nodes[i]);
    return root > 0 ? sum / root : 0;
}

void lro_sah_put_nodes(const LroSah *b, uint32_t node0, uint32_t ref0, uint32_t *nodes, uint32_t *aux)
{
    for (int i = 0; i < b->nnodes; i++) {
        const LroSahNode *n = &b->nodes[i];
        uint32_t a = n->count > 0 ? ref0 + (uint32_t)n->start : node0 + (uint32_t)n->left;
        uint32_t c = n->count > 0 ? (0x80000000u | (uint32_t)n->count) : node0 + (uint32_t)n->right;
        put_node(nodes + 8 * (size_t)i, n->lo, a, n->hi, c);
    }
    for (int i = 0; i < b->nrefs; i++) aux[i] = (uint32_t)b->refs[i];
}
