/* lro.h - the oracle: a plain-C, CPU-only restatement of little_renderer's path-tracing hot
 * path (BVH builders, traversal, sampler, BSDF, path tracer).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP implementation is compared
 * against. Only the tests, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library (little_renderer_amd/liblr_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line whose rule it restates. Floating point:
 * compiled with -ffp-contract=off, so every a*b+c rounds twice, exactly as written (the
 * reference's float build does the same on x86-64: no FMA without -march, Makefile:22-33).
 *
 * Layouts shared with the device (the reference's GPU layouts, trace_gpu.c:860-911):
 *   box   8 floats: lo.x lo.y lo.z 0 hi.x hi.y hi.z 0          (2 x vec4, build_queue :587)
 *   node  8 words : lo.xyz (f32), a (u32), hi.xyz (f32), b (u32) (trace.wgsl:2)
 *           inner: a, b = absolute child node indices
 *           leaf : a = first aux slot, b = 0x80000000 | count
 *   tri   12 words: q.xyz n0 | u.xyz n1 | v.xyz (mat<<26 | n2)   (put_tris :898-911)
 */
#ifndef LRO_H
#define LRO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers (engine/base/rng.h) ---- */
typedef struct { uint64_t state, inc; } LroRng;
uint32_t lro_rng_u32(LroRng *r);                                  /* rng.h:13-20 */
void lro_rng_seed(LroRng *r, uint64_t seed, uint64_t stream);     /* rng.h:22-28 */
float lro_rng_float(LroRng *r);                                   /* rng.h:31-37 (float build) */
double lro_rng_double(LroRng *r);                                 /* rng.h:31-37 (double build) */

/* ---- BVH builds (trace_build.wgsl, trace_ploc.wgsl, trace_gpu.c:545-823) ---- */
#define LRO_LEAF 0x80000000u
/* triangle_bounds (scene.c:65-74) of n packed triangles -> boxes (8 floats each) */
void lro_tri_boxes(const float *tris, int n, float *boxes);
/* the union of n boxes (lo, hi as 3 floats) */
void lro_union(const float *boxes, int n, float lo[3], float hi[3]);
/* build_queue's 1 / extent (trace_gpu.c:607-611): (float)(1.0 / (double)(hi - lo)), 0 if flat */
void lro_inv_extent(const float lo[3], const float hi[3], float ext[3]);
/* morton (trace_build.wgsl:16-33): the 30-bit code of each box's centre */
void lro_morton(const float *boxes, int n, const float lo[3], const float ext[3], uint32_t *codes);
/* the sorted order of the keys (code, item): bsort's result (trace_build.wgsl:35-50) */
void lro_sort(const uint32_t *codes, int n, uint32_t *sorted_codes, uint32_t *sorted_items);
/* LBVH: hier + leaves + up (trace_build.wgsl:52-122) over sorted keys. nodes: 2n-1 records
 * (node0 is the first), aux: n words (ref0 the first). Returns the tree's depth. */
int lro_lbvh(const uint32_t *sorted_codes, const uint32_t *sorted_items, const float *boxes, int n,
             uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux);
/* the same tree made bottom-up (Apetrei 2014): an independent construction of the binary radix
 * tree, with Karras's node numbering; must equal lro_lbvh to the bit */
int lro_lbvh_bottom_up(const uint32_t *sorted_codes, const uint32_t *sorted_items, const float *boxes, int n,
                       uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux);
/* PLOC (trace_ploc.wgsl:1-301) with radius 16 over the Morton-ordered leaves. Returns the
 * number of rounds; *depth the tree's depth; *live (optional) the clusters summed over the
 * rounds. n <= 2 builds an LBVH, as build_queue does. */
int lro_ploc(const uint32_t *sorted_codes, const uint32_t *sorted_items, const float *boxes, int n,
             uint32_t node0, uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux, int *depth,
             uint64_t *live);
/* check_tree (trace_gpu.c:768-805): 1 if nodes are one tree over exactly n leaves; its cost
 * (sum of areas / root's, in double) and depth */
int lro_check_tree(const uint32_t *nodes, int n, uint32_t node0, double *cost, int *depth);
/* the whole build from triangles: boxes, bounds, morton, sort, then LBVH (ploc = 0) or PLOC.
 * keys_out (optional): the sorted codes and items, 2n words. */
int lro_build_tris(const float *tris, int n, int ploc, uint32_t node0, uint32_t ref0, uint32_t ref_add,
                   uint32_t *nodes, uint32_t *aux, uint32_t *keys_out);

/* ---- the CPU's binned-SAH BVH (engine/rt/bvh.c), float build ---- */
typedef struct {
    float lo[3], hi[3];
    int left, right, start, count;           /* bvh.h:20-24 */
} LroSahNode;
typedef struct {
    float b[6][4];
    int child[4], count[4];                  /* bvh.h:31-35 */
} LroWideNode;
typedef struct {
    LroSahNode *nodes; int nnodes;
    int *refs; int nrefs;
    LroWideNode *wide; int nwide;
    int leaves, depth;
} LroSah;
/* bvh_build_boxes (bvh.c:405-461) over n boxes, max_leaf, bins; then bvh_make_wide if wide */
LroSah *lro_sah_build(const float *boxes, int n, int max_leaf, int bins, int wide);
void lro_sah_free(LroSah *b);
/* node count, ref count, depth and wide-node count of a built tree */
void lro_sah_info(const LroSah *b, int *nnodes, int *nrefs, int *depth, int *nwide);
double lro_sah_cost(const LroSah *b);        /* bvh_cost (bvh.c:464-469) */
/* the SAH tree in the GPU's binary layout (put_nodes, trace_gpu.c:883-895): nnodes records */
void lro_sah_put_nodes(const LroSah *b, uint32_t node0, uint32_t ref0, uint32_t *nodes, uint32_t *aux);

/* ---- traversal ---- */
/* one ray: o, d, tmin, tmax (8 floats: o.xyz tmin d.xyz tmax) -> hit (t, id, u, v) with
 * id = 0xffffffff for a miss. any = 1 stops at the first hit. */
typedef struct { float t; uint32_t id; float u, v; } LroHit;
typedef struct { uint64_t nodes, tris; } LroCount;
/* trace.wgsl traverse (:270-392) over a binary tree in the GPU layout (one mesh of triangles,
 * shape id = the aux entry), with an unbounded stack. counts: node records fetched and
 * triangles tested, summed over the rays (optional). */
void lro_trace_gpu_layout(const uint32_t *nodes, uint32_t root, const uint32_t *aux, const float *tris,
                          const float *rays, int nrays, int any, LroHit *hits, LroCount *count);
/* the CPU's walk (bvh.c:541-663) over the SAH tree, binary or 4-wide as built */
void lro_trace_sah(const LroSah *b, const float *tris, const float *rays, int nrays, int any, LroHit *hits);
/* every triangle tested, in id order (the brute force); second: the second-nearest t */
void lro_trace_brute(const float *tris, int ntris, const float *rays, int nrays, int any, LroHit *hits, float *second);

#ifdef __cplusplus
}
#endif
#endif
