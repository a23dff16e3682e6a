// lr_engine.h - the path-tracing engine's device state: many jobs (each its own resident
// scene, camera, image and sample range) share one pool of path slots that the wavefront
// kernels keep full.
//
// One iteration is four kernels: regen (start new paths at the end of the ray queue), extend
// (closest hits of the ray queue), shade (emission, next-event shadow ray, BSDF sample, Russian
// roulette), shadow (any hit before the lamp). A path's state lives beside its ray, at its
// position in the ray queue (two queues, one per iteration parity): shade reads entry i's state
// and writes a surviving path's state where its next ray is appended, so every state access is
// coalesced. A path that ends frees one place of the pool; regen fills free places from the
// global path index space, in which each job owns chunks of [0, w*h*spp) (sample-major:
// q = sample * w*h + pixel), each chunk an issue record mapping a range of global path indices
// to (job, first q). The same kernels serve
// the reference's one-job-at-a-time shape (a grid per pass, passes in order) and the batched
// form (persistent work-groups, one hipGraph of several iterations replayed until the work is
// done), so a job's image is the same to the bit in either, alone or among others: each
// sample depends only on (job seed, pixel, sample index) and is added to its pixel as a
// fixed-point integer (2^-28 units), whose sum does not depend on order.
#pragma once

#include <vector>

#include "lr_internal.h"

#define LR_SLOT_BITS 22                   // pool places: at most 2^22 (shadow ids: job << 22 | next queue position)
#define LR_SLOT_MASK ((1u << LR_SLOT_BITS) - 1u)
#define LR_JOBS 1024                      // job descriptors held on the device at once
#define LR_ISSUE 4096                     // issue records (a ring, power of two)
#define LR_FIX_SHIFT 28                   // a sample's radiance in units of 2^-28
#define LR_FIX_CAP 16777216.0f            // and saturated at 2^24 (a pixel can hold 2^11 such)

// one job on the device (written by a kernel with the descriptor as its argument, so the
// host's copy is never read late)
struct JobDev {
    const float4 *nodes;      // 4-wide nodes (walk4); a two-level scene: its top level, 4-wide (walk4_tl)
    const float4 *bnodes;     // a two-level scene's top level as binary records (k_child_boxes; not walked now)
    const uint32_t *aux;
    const float4 *tris;
    const float4 *ltris;      // the triangles again, in leaf order: ltris[3 e..] for aux entry e (the walks)
    const float4 *prims;
    const float4 *mats;       // 3 float4 per material: albedo.rgb type | emission.rgb ior | metallic roughness transmission 0
    long long *accum;         // 3 x int64 per pixel, fixed point
    unsigned long long total; // w * h * spp
    float cam[12];
    uint32_t root, nprims;
    int W, H, npix, light;
    int max_depth, mis;
    uint32_t s0, seed_lo, seed_hi;
    float clamp, eps, margin;
    // the sky (none: env_w = 0)
    const float *env_rgb, *env_marg, *env_cond, *env_prob;
    int env_w, env_h;
    float env_scale, env_rot, env_rot_c, env_rot_s;
    // a two-level scene: nodes/bnodes/aux above are the top level over the instances (aux:
    // instance ids); the meshes' binary records, 4-wide nodes and aux (global triangle ids into
    // tris; ltris then holds the meshes' triangles in maux order); per instance 4 float4: world
    // -> object rows, then (kind, material, mesh root record, mesh root wide slot) as u32
    const float4 *mnodes;
    const float4 *mnodes4;
    const uint32_t *maux;
    const float4 *inst;
    int two_level;
    // the glowing shapes sampled as lights (Scene.emitters): 5 float4 each, (centre, ball 1/0 |
    // half axis 0, radius | half axis 1, area | half axis 2, 0 | radiance, 0); each instance's
    // slot among them (-1: none)
    const float4 *emit;
    const int *inst_emit;
    int nemit;
    // a job with the denoiser's inputs (else both null): per pixel its guide, 2 float4 (the
    // first diffuse surface's normal, distance | the albedo on the way to it, the kind: 0 a
    // surface, 1 the sky or a lamp), and the moments of its samples' irradiance luminance (two
    // int64: the sum in 2^-28 units, the sum of squares in 2^-20 units)
    const float4 *guide;
    long long *moments;
    // a job with a target error, in its first 4 passes: each pass's luminance summed (fixed point,
    // 2^-28) into the progress block, for the host to choose the pass count from
    int pilot;
    // light tracing on: the paths the light tracer covers (the camera saw a rough surface first,
    // and the lamp is reached right after a mirror or glass) leave the lamp out
    int skip_caustics;
    // light paths filling a photon cache (lrh_photons_trace): its table and finest cell's size
    unsigned long long *pc;
    float pc_cell;
};
// the photon cache: LR_PC_SLOTS slots of 4 x 64 bits (rgb power sums in 2^-8 units; the photons
// in 1/16 units, 32 bits, and the cell's 32-bit tag, 0 = free, in the fourth word)
#define LR_PC_SLOTS (1u << 20)
#define LR_PC_SCALE 256.0f
#define LR_PC_UNIT 16u
#define LR_PILOT_PASSES 4
#define LR_LIGHT_CHUNK (1 << 22)          // light paths a launch pair (the queue's entries: 256 MB)
#define LR_MOM2_SHIFT 20                  // the squares' fixed-point units: 2^-20
#define LR_MOM2_CAP 1099511627776.0f      // a square saturated at 2^40
#define LR_MAX_EMITTERS 16

// the compressed 4-wide nodes' leaf references hold a first aux entry in 24 bits (all ones is
// kept free: a two-level walk's leaf entry then never equals its EXIT entry)
#define LR_WIDE_MAX_SHAPES ((1 << 24) - 1)

// instance kinds on the device: a mesh index, or these (trace.wgsl:188-190)
#define LR_UNIT_SPHERE 0xfffffff0u
#define LR_UNIT_BOX 0xfffffff1u
// stack entries of the two-level walks: a top-level node k (< 2^30), INST | instance, BLAS | node
// (and, in the 4-wide walk, a mesh leaf: 0xc0000000 | (count - 1) << 24 | first maux entry)
#define LR_WALK_INST 0x40000000u
#define LR_WALK_BLAS 0x80000000u
#define LR_WALK_EXIT 0xffffffffu

enum { K_REGEN = 0, K_EXTEND = 1, K_SHADE = 2, K_SHADOW = 3 };

// the engine's counters (device memory, followed by each job slot's progress)
struct EngCtl {
    unsigned long long head, tail;          // free places of the pool: tail - head
    unsigned long long next, issue_end;     // the next path index to start; the end of those issued
    unsigned long long iq_lo, iq_hi;        // the issue records in use
    unsigned long long rays_ext, rays_shadow;
    unsigned long long iter;                // iterations run (its parity picks the ray queue)
    unsigned long long claimed;             // paths the running iteration's regen started
    unsigned int count[4];                  // ray queues [0], [1]; shadow queue [2]
    unsigned int tiles[4];                  // each kernel's tile counter (work-groups claim tiles)
    unsigned int regen_done;                // regen's work-groups that have finished
    unsigned int spar;                      // the parity of the queue the shade kernel appended to
    unsigned long long pad[3];
};

struct EngProgress {
    EngCtl ctl;
    unsigned long long jret[LR_JOBS];       // paths retired per job slot
    unsigned long long jdone[LR_JOBS];      // device clock when the slot's last path retired
    long long jpass[LR_JOBS][LR_PILOT_PASSES];   // a pilot job's first passes: their luminance sums
};

// everything a kernel needs, passed by value
struct EngArgs {
    EngCtl *ctl;
    const JobDev *jobs;
    unsigned long long *jret, *jdone;
    long long (*jpass)[LR_PILOT_PASSES];
    const unsigned long long *iq_start;     // issue record: the first global path index it covers,
    const uint32_t *iq_job;                 // its job slot,
    const uint32_t *iq_q0;                  // and the index in the job of its first path
    // a path's state, by ray queue (parity) and queue position
    float4 *thr[2];                         // throughput.rgb, the last bounce's pdf
    float4 *rad[2];                         // radiance.rgb, the emission state
    uint64_t *rng[2];                       // PCG32 state (stream: the pixel)
    uint2 *meta[2];                         // q (index in the job), depth
    float4 *qray[2];                        // o.xyz tmin | d.xyz tmax
    uint32_t *qid[2];                       // job << 22
    float4 *hits;                           // t, id, b1, b2 (by queue position)
    float4 *sray;
    uint32_t *sid;                          // job << 22 | the path's next queue position
    float4 *scon;                           // contribution.rgb, 1 if the path ends after it
    float4 *srad;                           // the path's radiance and emission state before it
    uint32_t *sq;                           // the path's q
    // a vertex with more than one light (lamp, sky, glowing shapes) has more shadow rays: the
    // first in the entry above, the others beside shade's input position, and sx the count of
    // those (bits 30-31) and that position
    float4 *xsray[2];
    float4 *xscon[2];
    uint32_t *sx;
    uint32_t *spill;                        // walk stacks past the LDS part, by thread
    unsigned long long spill_stride;        // threads of the launch
    int spill_per;
    uint32_t *overflow;
    const float *tab;                       // BSDF tables
    const uint16_t *mask;                   // blue-noise ranks
    unsigned long long *counts;             // node records and shapes walked (counting extend only)
    unsigned long long *ostat;              // shade's tiles: look-back status of its ordered appends
    int mult;                               // a tile is mult work-groups' widths of entries
    int refill;                             // k_extend_rf: idle lanes of a wave that take new rays
    int refill_shadow;                      // k_shadow_rf: the same
};

typedef void (*LrEngFn)(dim3 grid, hipStream_t s, const EngArgs &A);
// the denoiser's guide of one job (k_guide, 256 threads, a 16-entry LDS stack)
typedef void (*LrGuideFn)(dim3 grid, hipStream_t s, const JobDev &J, float4 *guide, uint32_t *spill, size_t stride,
                          int spill_per, const float *tab, uint32_t *overflow);
// light paths g0 .. g0+n-1 of a job (path g: pass g / nlp) for caustics (k_light_first, then
// k_light_rest over the queue of those that go on; 256 threads, a 16-entry LDS stack; n at most
// the engine's threads); the rays they walk added to cnt[0] (closest hit) and cnt[1] (towards the
// camera), and in the counting form their node records and shapes to cnt[2] and cnt[3]
typedef void (*LrLightFn)(hipStream_t s, const JobDev &J, uint32_t nlp, uint64_t g0, uint32_t n, uint32_t *spill,
                          int spill_per, const float *tab, uint32_t *overflow, float4 *queue, uint32_t *qcount,
                          unsigned long long *cnt);

// the kernels, by work-group size (64, 128, 256: index 0..2), LDS stack (16, 24, 32) and, for the
// walks, whether they take two-level scenes too (0: one-level scenes only; the two-level walks'
// registers would lower the one-level walks' occupancy); shade and shadow also by whether a
// vertex may have more than one light (1: a lamp and a sky, glowing shapes; 0: one shadow ray,
// in fewer registers)
struct LrEngTable {
    LrEngFn regen[3], shade[3][2];
    LrEngFn extend[3][3][2], extend_count[3][3][2], shadow[3][3][2][2], shadow_count[3][3][2][2];
    // [block][stack][counting]: the walks with lanes refilled (single-level scenes; the shadow
    // walk's with one shadow ray a vertex)
    LrEngFn extend_rf[3][3][2], shadow_rf[3][3][2];
    LrGuideFn guide[2];   // by whether it takes two-level scenes
    LrLightFn light[2];   // by counting (one-level scenes: only they have a lamp)
};
void lr_eng_fill_b64(LrEngTable &t);
void lr_eng_fill_b128(LrEngTable &t);
void lr_eng_fill_b256(LrEngTable &t);

static inline int lr_blk_index(int b) { return b == 64 ? 0 : b == 128 ? 1 : b == 256 ? 2 : -1; }
static inline int lr_stk_index(int s) { return s == 16 ? 0 : s == 24 ? 1 : s == 32 ? 2 : -1; }

// a scene resident on the device (lrh_pt_scene)
struct LrhPtScene {
    LrhContext *ctx;
    float4 *prims = nullptr, *tris = nullptr, *mats = nullptr, *boxes = nullptr;
    uint32_t *nodes = nullptr, *aux = nullptr;
    float4 *wnodes = nullptr;                 // a two-level scene's top level as binary records (k_child_boxes)
    float4 *w4nodes = nullptr;                // the same tree, 4-wide (walk4)
    float4 *ltris = nullptr;                  // the triangles in leaf order (zeros at primitives' entries)
    // the sky
    float *env_rgb = nullptr, *env_marg = nullptr, *env_cond = nullptr, *env_prob = nullptr;
    int env_w = 0, env_h = 0;
    float env_scale = 1.0f, env_rot = 0.0f;
    // a two-level scene (lrh_pt_scene_instanced): meshes, instances, the top level
    int two_level = 0, nmeshes = 0, ninst = 0, mesh_depth = 0, tlas_depth = 0;
    int top_builder = 0;                      // the top level's last builder (a mesh's update builds it again so)
    uint32_t *tparent = nullptr, *tarrive = nullptr;   // the top level's records' parents; refit arrivals
    double *tcost_dev = nullptr;              // a cost computed on the device
    double tcost_build = 0.0;                 // the top level's cost at its last build (the refit rule's)
    float4 *mtris = nullptr, *mboxes = nullptr, *mwnodes = nullptr, *inst = nullptr, *ifwd = nullptr;
    float4 *mw4nodes = nullptr, *tw4nodes = nullptr;   // the meshes' and the top level's 4-wide nodes
    uint32_t *mroot4 = nullptr;                        // each mesh's root wide slot
    uint32_t *mnodes = nullptr, *maux = nullptr, *mroot = nullptr;
    float4 *tboxes = nullptr;
    uint32_t *tnodes = nullptr, *taux = nullptr;
    // host copies of the meshes' layout: first triangle (nmeshes + 1 of them), root record,
    // root wide slot and tree depth of each mesh
    std::vector<int> mfirst, mdepth;
    std::vector<uint32_t> mroot_h, mroot4_h;
    float *tab = nullptr;
    uint16_t *mask = nullptr;
    // the glowing shapes (an instanced scene's), from host copies of the instances and materials
    float4 *emit = nullptr;
    int *inst_emit = nullptr;
    int nemit = 0;
    std::vector<LrhInstance> hinst;
    std::vector<LrhMaterial> hmats;
    int nprims = 0, ntris = 0, nmats = 0, light = -1, depth = 1;
    float eps = 1e-4f, margin = 1e-3f;
    size_t bytes = 0;                       // device memory it holds
};

// a scene whose vertices may need more than one shadow ray
static inline bool lr_scene_multi_light(const LrhPtScene *s)
{
    return (s->light >= 0 && s->env_w > 0) || s->nemit > 0;
}

// a job's block in the engine's image arena: its image (24 bytes a pixel) and, when it is to be
// denoised, its guide (32) and moments (16)
static inline size_t lr_engine_job_bytes(const LrhJob &j)
{
    const size_t npix = (size_t)j.w * (size_t)j.h;
    return lr_align(npix * 24) + (j.denoise ? lr_align(npix * 32) + lr_align(npix * 16) : 0);
}

// the engine with its image arena made at once, or chunk by chunk as images need it up to
// image_bytes (lazy); a lazy arena's cap, the chunks' bytes, and its chunks holding no image freed
int lr_engine_create(LrhContext *c, int slots_log2, int max_tree_depth, size_t image_bytes, bool lazy, LrhEngine **out);
void lr_engine_arena_cap(LrhEngine *e, size_t cap);
size_t lr_engine_arena_held(const LrhEngine *e);
void lr_engine_arena_trim(LrhEngine *e);

// a job descriptor's scene part (the shapes, trees, materials, sky) from a resident scene
void lr_job_scene(JobDev &jd, const LrhPtScene *s);
