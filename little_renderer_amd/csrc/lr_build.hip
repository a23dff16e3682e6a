// lr_build.hip - the BVH builders on gfx950: LBVH (Karras 2012 numbering, built bottom-up in
// one pass after Apetrei 2014) and PLOC (Meister & Bittner 2018, radius 16), over Morton
// codes sorted by an LSD radix sort with decoupled look-back (one kernel per 8-bit digit).
//
// The rules are the reference's (engine/gpu/shaders/trace/trace_build.wgsl,
// trace_ploc.wgsl); the kernels are not: the reference's 210 bitonic steps become 4 radix
// passes, its 54 refit passes one bottom-up pass with per-node arrival counters, and each
// PLOC round (4 dispatches there) one kernel whose cluster windows sit in LDS.
#include <stdarg.h>
#include <stddef.h>
#include <algorithm>
#include <mutex>
#include <stdio.h>

#include "lr_device.h"
#include "lr_internal.h"

namespace {

// Wait until this thread's memory operations are done. Device-scope atomics and stores are
// coherent across the XCDs once done, so a hand-off needs this wait before its flag, not a
// __threadfence (on gfx950 a write-back of the whole L2, here once per wave)
__device__ __forceinline__ void wait_mem()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}


constexpr int BLOCK = 256;
constexpr int SORT_BLOCK = 1024;                // 16 waves per tile: enough in flight per CU
constexpr int SORT_ITEMS = 4;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;   // 4096 keys per tile
constexpr int SORT_WAVES = SORT_BLOCK / 64;
constexpr int SORT_PASSES = 4;                  // 30-bit codes: bits 0-7, 8-15, 16-23, 24-29
constexpr int HIST_COPIES = 8;                  // one global histogram per XCD (less contention)
constexpr int PLOC_R = 16;
constexpr int PLOC_PER = 2;                     // positions a thread takes in a PLOC round
constexpr int PLOC_TILE = 256 * PLOC_PER;        // positions a work-group takes (one claim)
constexpr int PLOC_WIN = PLOC_TILE + 4 * PLOC_R;   // boxes of positions base-2R .. base+PLOC_TILE+2R
constexpr int PLOC_NN = PLOC_TILE + 2 * PLOC_R;    // nearest of positions base-R .. base+PLOC_TILE+R
constexpr int PLOC_MAX_ROUNDS = 128;
constexpr int BOUNDS_BLOCKS = 512;

constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_PREFIX = 2u << 30;
constexpr int LOOK_BACK = 8;                     // predecessor tiles read at once in a sort pass's look-back
constexpr uint32_t VALUE_MASK = (1u << 30) - 1u;

// ---- block-wide scans (256 threads, wave64) ------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// exclusive scan of one value per thread over a block of NW waves; *total gets the sum
template <int NW>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total, uint32_t *s_wsum)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        uint32_t s = s_wsum[k];
        off += k < w ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// ---- triangle boxes (scene.c:65-74) --------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_tri_boxes(const float4 *__restrict__ tris, int n, float4 *__restrict__ boxes)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 q = tris[3 * (size_t)i], u = tris[3 * (size_t)i + 1], v = tris[3 * (size_t)i + 2];
    const float qq[3] = {q.x, q.y, q.z}, uu[3] = {u.x, u.y, u.z}, vv[3] = {v.x, v.y, v.z};
    float lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float a = qq[k] + uu[k], b = qq[k] + vv[k];
        lo[k] = rfmin(rfmin(qq[k], a), b);
        hi[k] = rfmax(rfmax(qq[k], a), b);
    }
    boxes[2 * (size_t)i] = make_float4(lo[0] - 1e-3f, lo[1] - 1e-3f, lo[2] - 1e-3f, 0.0f);
    boxes[2 * (size_t)i + 1] = make_float4(hi[0] + 1e-3f, hi[1] + 1e-3f, hi[2] + 1e-3f, 0.0f);
}

// ---- the bounds of the boxes, and the Morton parameters ---------------------------------------
__device__ __forceinline__ void min6(float *a, const float *b)
{
#pragma unroll
    for (int k = 0; k < 3; k++) { a[k] = rfmin(a[k], b[k]); a[3 + k] = rfmax(a[3 + k], b[3 + k]); }
}

__device__ void block_min6(float *v, float (*s)[6])
{
    // within the wave, then across the 4 waves
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float t[6];
#pragma unroll
        for (int k = 0; k < 6; k++) t[k] = __shfl_xor(v[k], o, 64);
        min6(v, t);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 6; k++) s[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 1; q < BLOCK / 64; q++) min6(v, s[q]);
}

__global__ __launch_bounds__(BLOCK) void k_bounds(const float4 *__restrict__ boxes, int n, float *partial,
                                                    uint32_t *ticket, LrBuildParamsDev *out)
{
    __shared__ float s[BLOCK / 64][6];
    __shared__ bool last;
    float v[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const float4 a = boxes[2 * (size_t)i], b = boxes[2 * (size_t)i + 1];
        const float t[6] = {a.x, a.y, a.z, b.x, b.y, b.z};
        min6(v, t);
    }
    block_min6(v, s);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; k++) st_dev((uint32_t *)&partial[6 * blockIdx.x + k], __float_as_uint(v[k]));
        // the ticket releases this block's partials and, for the last block, acquires the others'
        last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    float w[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += BLOCK) {
        float t[6];
        for (int k = 0; k < 6; k++) t[k] = __uint_as_float(ld_dev((const uint32_t *)&partial[6 * b + k]));
        min6(w, t);
    }
    __syncthreads();
    block_min6(w, s);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; k++) {
            const float e = w[3 + k] - w[k];
            out->lo[k] = w[k];
            out->hi[k] = w[3 + k];
            out->ext[k] = e > 0.0f ? (float)(1.0 / (double)e) : 0.0f;   // trace_gpu.c:609-611
        }
        out->lo[3] = out->ext[3] = out->hi[3] = 0.0f;
        *ticket = 0;
    }
}

// bounds given by the caller: only the parameters
__global__ void k_params(LrBuildParamsDev *out, float lx, float ly, float lz, float hx, float hy, float hz)
{
    const float lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
    for (int k = 0; k < 3; k++) {
        const float e = hi[k] - lo[k];
        out->lo[k] = lo[k];
        out->hi[k] = hi[k];
        out->ext[k] = e > 0.0f ? (float)(1.0 / (double)e) : 0.0f;
    }
    out->lo[3] = out->ext[3] = out->hi[3] = 0.0f;
}

// ---- Morton codes (trace_build.wgsl:16-33) and the sort's global histograms ----------------------
__device__ __forceinline__ uint32_t spread(uint32_t v)
{
    uint32_t x = v & 1023u;
    x = (x | (x << 16u)) & 0x030000ffu;
    x = (x | (x << 8u)) & 0x0300f00fu;
    x = (x | (x << 4u)) & 0x030c30c3u;
    x = (x | (x << 2u)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t quant(float c, float lo, float ext)
{
    return (uint32_t)(wmin(wmax((c - lo) * ext, 0.0f), 0.999f) * 1024.0f);
}

// one tile of 4096 items per block: codes (and the items 0..n-1 as the sort's values), the
// four digits' counts into the histograms; the last block to finish turns them into each
// digit's start (exclusive scans) and clears them for the next build. Also clears the memory
// the later kernels expect zeroed (look-back states, tile counters, arrival counters).
__global__ __launch_bounds__(SORT_BLOCK) void k_morton(const float4 *__restrict__ boxes, int n, const LrBuildParamsDev *prm,
                                                    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                    uint32_t *hist, uint32_t *ticket, uint32_t *bucket_base,
                                                    uint32_t *zero, size_t zero_words)
{
    __shared__ uint32_t h[SORT_PASSES][256];
    __shared__ uint32_t s_wsum[SORT_WAVES];
    __shared__ bool last;
    for (int i = threadIdx.x; i < SORT_PASSES * 256; i += SORT_BLOCK) (&h[0][0])[i] = 0;
    for (size_t i = blockIdx.x * (size_t)SORT_BLOCK + threadIdx.x; i < zero_words; i += (size_t)gridDim.x * SORT_BLOCK) zero[i] = 0;
    __syncthreads();
    const float lx = prm->lo[0], ly = prm->lo[1], lz = prm->lo[2];
    const float ex = prm->ext[0], ey = prm->ext[1], ez = prm->ext[2];
    const int base = blockIdx.x * SORT_TILE;
    // all the tile's boxes in flight at once, then the codes
    float4 a[SORT_ITEMS], b[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        const int i = base + r * SORT_BLOCK + threadIdx.x;
        if (i < n) {
            a[r] = boxes[2 * (size_t)i];
            b[r] = boxes[2 * (size_t)i + 1];
        }
    }
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        const int i = base + r * SORT_BLOCK + threadIdx.x;
        if (i >= n) continue;
        const float cx = (a[r].x + b[r].x) * 0.5f, cy = (a[r].y + b[r].y) * 0.5f, cz = (a[r].z + b[r].z) * 0.5f;
        const uint32_t code = (spread(quant(cx, lx, ex)) << 2u) | (spread(quant(cy, ly, ey)) << 1u) | spread(quant(cz, lz, ez));
        keys[i] = code;
        vals[i] = (uint32_t)i;
#pragma unroll
        for (int p = 0; p < SORT_PASSES; p++) atomicAdd(&h[p][(code >> (8 * p)) & 255u], 1u);
    }
    __syncthreads();
    uint32_t *mine = hist + (blockIdx.x % HIST_COPIES) * SORT_PASSES * 256;
    for (int i = threadIdx.x; i < SORT_PASSES * 256; i += SORT_BLOCK) {
        const uint32_t v = (&h[0][0])[i];
        if (v) atomicAdd(&mine[i], v);
    }
    // The ticket hands the histograms over. They move as device-scope atomics (coherent across
    // the XCDs once done), so the ticket needs every thread's adds done before it (wait_mem:
    // the wait, and compiler barriers on both sides; then the work-group barrier), and the last
    // work-group reads them with device-scope loads after it. An acquire-release ticket instead
    // writes back and invalidates the L2 once per work-group (about 3900 of them at 16M keys):
    // measured 1.68 -> 1.89 ms for a 16M build together with the bounds kernel's.
    wait_mem();
    __syncthreads();
    if (threadIdx.x == 0) last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    // the last block: thread t sums digit t % 256 of pass t / 256 over the copies (all the
    // copies' loads in flight together), clears them, and the passes are scanned side by side
    // (four waves per pass)
    static_assert(SORT_BLOCK == SORT_PASSES * 256, "one thread per (pass, digit)");
    uint32_t c8[HIST_COPIES];
#pragma unroll
    for (int c = 0; c < HIST_COPIES; c++) c8[c] = ld_dev(&hist[c * SORT_PASSES * 256 + threadIdx.x]);
    uint32_t v = 0;
#pragma unroll
    for (int c = 0; c < HIST_COPIES; c++) {
        v += c8[c];
        st_dev(&hist[c * SORT_PASSES * 256 + threadIdx.x], 0u);
    }
    const int w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if ((threadIdx.x & 63) == 63) s_wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int k = w & ~3; k < w; k++) off += s_wsum[k];
    bucket_base[threadIdx.x] = off + inc - v;
    if (threadIdx.x == 0) *ticket = 0;
}

// ---- one radix-sort pass: 8 bits, stable, decoupled look-back (onesweep) -------------------------
// Each wave ranks its 1024 keys (16 rounds of 64, wave-contiguous, so the rank order is the
// input order) with a 64-lane match over the digit's 8 bits; per-wave digit counts give each
// key its place in the tile; the tile's place in each digit's run comes from the tiles before
// it (look-back over per-tile, per-digit states). The tile is reordered in LDS first, so that
// the stores to each digit's run are contiguous.
__global__ __launch_bounds__(SORT_BLOCK) void k_sort_pass(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                       uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, int n,
                                                       int shift, const uint32_t *__restrict__ bucket_base,
                                                       uint32_t *status, uint32_t *tile_ctr)
{
    __shared__ uint32_t s_tile;
    __shared__ uint32_t wc[SORT_WAVES][256];
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_glob[256];
    __shared__ uint32_t s_wsum[SORT_WAVES];
    __shared__ uint32_t s_k[SORT_TILE];
    __shared__ uint32_t s_v[SORT_TILE];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (int i = tid; i < SORT_WAVES * 256; i += SORT_BLOCK) (&wc[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const int base = (int)tile * SORT_TILE;
    uint32_t k[SORT_ITEMS], v[SORT_ITEMS], rank[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        const int i = base + w * (64 * SORT_ITEMS) + r * 64 + lane;
        const bool ok = i < n;
        k[r] = ok ? kin[i] : 0u;
        v[r] = ok ? vin[i] : 0u;
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        const int i = base + w * (64 * SORT_ITEMS) + r * 64 + lane;
        const bool ok = i < n;
        const uint32_t d = (k[r] >> shift) & 255u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t before = ok ? wc[w][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(peers & lt);
        __builtin_amdgcn_wave_barrier();
        if (ok && (peers & lt) == 0) wc[w][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // threads 0..255: digit tid
    const bool dig = tid < 256;
    uint32_t tot = 0;
    if (dig) {
        for (int q = 0; q < SORT_WAVES; q++) { const uint32_t c = wc[q][tid]; wc[q][tid] = tot; tot += c; }
        if (tile == 0) st_dev(&status[tid], FLAG_PREFIX | tot);
        else st_dev(&status[tile * 256 + tid], FLAG_AGG | tot);
    }
    uint32_t all;
    const uint32_t start = block_excl_scan<SORT_WAVES>(tot, &all, s_wsum);
    uint32_t excl = 0;
    if (dig && tile > 0) {
        // LOOK_BACK tiles at a time (their loads in flight together): the nearest prefix among
        // them ends the walk; a tile not yet posted means the same window again
        int t = (int)tile - 1;
        for (;;) {
            uint32_t sv[LOOK_BACK];
#pragma unroll
            for (int k = 0; k < LOOK_BACK; k++) sv[k] = t - k >= 0 ? ld_dev(&status[(t - k) * 256 + tid]) : FLAG_PREFIX;
            bool wait = false, done = false;
            uint32_t add = 0;
#pragma unroll
            for (int k = 0; k < LOOK_BACK; k++) {
                if (done || wait) continue;
                const uint32_t f = sv[k] & ~VALUE_MASK;
                if (f == 0) wait = true;
                else {
                    add += sv[k] & VALUE_MASK;
                    done = f == FLAG_PREFIX;
                }
            }
            if (wait) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += add;
            if (done) break;
            t -= LOOK_BACK;
        }
        st_dev(&status[tile * 256 + tid], FLAG_PREFIX | (excl + tot));
    }
    if (dig) {
        s_start[tid] = start;
        s_glob[tid] = bucket_base[tid] + excl - start;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        const int i = base + w * (64 * SORT_ITEMS) + r * 64 + lane;
        if (i < n) {
            const uint32_t d = (k[r] >> shift) & 255u;
            const uint32_t pos = s_start[d] + wc[w][d] + rank[r];
            s_k[pos] = k[r];
            s_v[pos] = v[r];
        }
    }
    __syncthreads();
    const int cnt = n - base < SORT_TILE ? n - base : SORT_TILE;
    for (int s = tid; s < cnt; s += SORT_BLOCK) {
        const uint32_t key = s_k[s];
        const uint32_t g = s_glob[(key >> shift) & 255u] + (uint32_t)s;
        kout[g] = key;
        vout[g] = s_v[s];
    }
}

// ---- LBVH, bottom-up in one pass -----------------------------------------------------------------
// delta of adjacent keys k, k+1 (trace_build.wgsl:52-58)
__device__ __forceinline__ int dadj(const uint32_t *codes, int k)
{
    const uint32_t a = codes[k], b = codes[k + 1];
    return a == b ? 32 + __clz((uint32_t)k ^ (uint32_t)(k + 1)) : __clz(a ^ b);
}

// is the node over keys [l, r] (not the root) its parent's left child? (its parent splits at r)
__device__ __forceinline__ bool is_left(const uint32_t *codes, int n, int l, int r)
{
    return l == 0 || (r != n - 1 && dadj(codes, r) > dadj(codes, l - 1));
}

__device__ __forceinline__ void put_node(LrNode *p, float lx, float ly, float lz, uint32_t a, float hx, float hy, float hz, uint32_t b)
{
    float4 *q = reinterpret_cast<float4 *>(p);
    q[0] = make_float4(lx, ly, lz, __uint_as_float(a));
    q[1] = make_float4(hx, hy, hz, __uint_as_float(b));
}

// a node record written through to the coherence point (sc1 stores), so that another XCD's
// thread reading it with ld_dev64 sees it without a cache-wide write-back
__device__ __forceinline__ void put_node_dev(LrNode *p, float lx, float ly, float lz, uint32_t a, float hx, float hy, float hz,
                                             uint32_t b)
{
    uint64_t *q = reinterpret_cast<uint64_t *>(p);
    st_dev64(q + 0, ((uint64_t)__float_as_uint(ly) << 32) | __float_as_uint(lx));
    st_dev64(q + 1, ((uint64_t)a << 32) | __float_as_uint(lz));
    st_dev64(q + 2, ((uint64_t)__float_as_uint(hy) << 32) | __float_as_uint(hx));
    st_dev64(q + 3, ((uint64_t)b << 32) | __float_as_uint(hz));
}

// The same through a buffer resource: a record as two 16-byte stores (and loads), sc1 (device
// scope, through to the coherence point) as st_dev64 / ld_dev64 but half the transactions. The
// resource covers 2 GiB from the records' base (the host takes this path only below that).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int CPOL_SC1 = 16;   // the sc1 cache-policy bit of gfx950's memory instructions
__device__ __forceinline__ __amdgpu_buffer_rsrc_t node_rsrc(LrNode *base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ void put_node_buf(__amdgpu_buffer_rsrc_t rs, int idx, float lx, float ly, float lz, uint32_t a,
                                             float hx, float hy, float hz, uint32_t b)
{
    const u32x4 lo = {__float_as_uint(lx), __float_as_uint(ly), __float_as_uint(lz), a};
    const u32x4 hi = {__float_as_uint(hx), __float_as_uint(hy), __float_as_uint(hz), b};
    __builtin_amdgcn_raw_buffer_store_b128(lo, rs, idx * 32, 0, CPOL_SC1);
    __builtin_amdgcn_raw_buffer_store_b128(hi, rs, idx * 32 + 16, 0, CPOL_SC1);
}

// min-scans over a group of NT threads (NT / 64 waves): forward (the minimum of this and every
// earlier thread's value) and backward (this and every later thread's)
template <int NT>
__device__ __forceinline__ int block_min_scan(int v, bool backward, int *s_w)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = backward ? __shfl_down(v, o, 64) : __shfl_up(v, o, 64);
        if (backward ? lane + o < 64 : lane >= o) v = t < v ? t : v;
    }
    if (lane == (backward ? 0 : 63)) s_w[w] = v;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NT / 64; k++) {
        const int t = s_w[k];
        if (backward ? k > w : k < w) v = t < v ? t : v;
    }
    __syncthreads();
    return v;
}

// The LBVH's nodes bottom-up (leaves, trace_build.wgsl:98-107; up, :109-122): from each
// leaf a thread climbs; at each parent the first of the two children to arrive stops, the
// second joins their boxes (min / max of left, right), writes the parent with Karras's
// numbering (internal i at node0 + i, leaf j at node0 + n - 1 + j) and goes on.
//
// Two kernels. k_lbvh: a work-group's LBVH_GROUP leaves climb through the nodes whose key ranges lie
// inside the group's: the two children's threads are then both in the group, and hand their
// boxes over in LDS (a slot per side, an LDS arrival counter). Whether a node's parent lies
// inside is decided from the adjacent-key deltas d(k) alone: a left child [l, p]'s parent ends
// at the first k > p with d(k) < d(p), a right child [p + 1, r]'s starts after the last k < p
// with d(k) < d(p); a min-scan of the group's deltas each way answers both with one
// comparison. A node whose parent lies outside (the root of a largest subtree inside the
// group: at most LBVH_GROUP of them, about 10 on random keys) is written and listed in the group's
// slots. k_lbvh_top then climbs from the listed nodes, one wave per group's list, through
// device memory: the two children's threads are on any XCDs, whose L2s are not coherent, so a
// record is stored through to the coherence point (sc1), the stores are waited for (vmcnt 0),
// and only then does the thread mark its arrival, with its subtree's far end and height in the
// one 64-bit word it swaps in; the second thread reads the record back with sc1 loads (a leaf
// child's box from the input boxes). Kept apart, the few long climbs through device memory
// do not hold the groups of the first kernel, and their LDS, on the compute units.
constexpr int LBVH_GROUP = 512;      // leaves per work-group of k_lbvh (1024 measured 13 % slower: one per CU)
constexpr int LBVH_SLOTS = LBVH_GROUP;   // listed nodes per group, at most

__global__ __launch_bounds__(LBVH_GROUP) void k_lbvh(const uint32_t *__restrict__ codes, const uint32_t *__restrict__ items,
                                                const float4 *__restrict__ boxes, int n, uint32_t node0, uint32_t ref0,
                                                uint32_t ref_add, LrNode *__restrict__ nodes, uint32_t *__restrict__ aux,
                                                uint4 *__restrict__ esc, uint32_t *__restrict__ esc_count,
                                                uint32_t *depth_out)
{
    __shared__ int s_d[LBVH_GROUP + 1];        // d(B0 - 1 + i); -1 (below every delta) off the keys' ends
    __shared__ int s_pmin[LBVH_GROUP + 1];     // min of s_d[0 .. i]
    __shared__ int s_smin[LBVH_GROUP + 1];     // min of s_d[i .. LBVH_GROUP] (i >= 1)
    __shared__ float s_box[LBVH_GROUP][2][6];  // per parent split, per side: the arriving child's box
    __shared__ uint32_t s_end[LBVH_GROUP][2], s_h[LBVH_GROUP][2], s_arrive[LBVH_GROUP];
    __shared__ int s_w[LBVH_GROUP / 64];
    __shared__ uint32_t s_nesc;
    const int t = threadIdx.x;
    const int B0 = blockIdx.x * LBVH_GROUP;
    const int j = B0 + t;
    const bool live = j < n;
    float lx = 0.0f, ly = 0.0f, lz = 0.0f, hx = 0.0f, hy = 0.0f, hz = 0.0f;
    if (live) {
        const uint32_t item = items[j];
        const float4 blo = boxes[2 * (size_t)item], bhi = boxes[2 * (size_t)item + 1];
        lx = blo.x; ly = blo.y; lz = blo.z; hx = bhi.x; hy = bhi.y; hz = bhi.z;
        aux[j] = item + ref_add;
        if (n > 1) put_node(&nodes[n - 1 + j], lx, ly, lz, ref0 + (uint32_t)j, hx, hy, hz, 0x80000001u);
        else {
            put_node(&nodes[0], lx, ly, lz, ref0, hx, hy, hz, 0x80000001u);
            atomicMax(depth_out, 1u);
        }
    }
    if (n == 1) {   // (uniform: the one group returns before any barrier)
        if (t == 0) esc_count[0] = 0u;
        return;
    }
    const int dj = j < n - 1 ? dadj(codes, j) : -1;
    s_d[t + 1] = dj;
    if (t == 0) {
        s_d[0] = B0 > 0 ? dadj(codes, B0 - 1) : -1;
        s_nesc = 0u;
    }
    s_arrive[t] = 0u;
    __syncthreads();
    const int d0 = s_d[0];
    const int pm = block_min_scan<LBVH_GROUP>(dj, false, s_w);
    const int sm = block_min_scan<LBVH_GROUP>(dj, true, s_w);
    s_pmin[t + 1] = pm < d0 ? pm : d0;
    s_smin[t + 1] = sm;
    if (t == 0) s_pmin[0] = d0;
    __syncthreads();

    // d(k) = s_d[k - B0 + 1] for B0 - 1 <= k <= B0 + LBVH_GROUP - 1
    int l = j, r = j;      // the node reached: keys [l, r] (a leaf while l == r)
    int lc = 0, rc = 0;    // an inner node's children
    uint32_t h = 1;
    bool climbing = live;
    bool left = live && s_d[r - B0 + 1] > s_d[l - B0];
    while (climbing) {
        const int p = left ? r : l - 1;
        const int dp = s_d[p - B0 + 1];
        const bool inside = left ? (p <= B0 + LBVH_GROUP - 2 && s_smin[p - B0 + 2] < dp) : (p >= B0 && s_pmin[p - B0] < dp);
        if (l != r) put_node(&nodes[left ? r : l], lx, ly, lz, node0 + (uint32_t)lc, hx, hy, hz, node0 + (uint32_t)rc);
        if (!inside) {   // listed for k_lbvh_top (its record, or its leaf's, holds its box)
            const uint32_t k = atomicAdd(&s_nesc, 1u);
            esc[(size_t)blockIdx.x * LBVH_SLOTS + k] = make_uint4((uint32_t)l, (uint32_t)r, h, 0u);
            break;
        }
        const int sl = p - B0, side = left ? 0 : 1;
        float *bx = s_box[sl][side];
        bx[0] = lx; bx[1] = ly; bx[2] = lz; bx[3] = hx; bx[4] = hy; bx[5] = hz;
        s_end[sl][side] = (uint32_t)(left ? l : r);
        s_h[sl][side] = h;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (atomicAdd(&s_arrive[sl], 1u) == 0u) break;   // the sibling's thread goes on
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const float *ob = s_box[sl][side ^ 1];
        const int oend = (int)s_end[sl][side ^ 1];
        const uint32_t sh = s_h[sl][side ^ 1];
        const int L = left ? l : oend, R = left ? oend : r;
        if (left) {
            lx = wmin(lx, ob[0]); ly = wmin(ly, ob[1]); lz = wmin(lz, ob[2]);
            hx = wmax(hx, ob[3]); hy = wmax(hy, ob[4]); hz = wmax(hz, ob[5]);
        } else {
            lx = wmin(ob[0], lx); ly = wmin(ob[1], ly); lz = wmin(ob[2], lz);
            hx = wmax(ob[3], hx); hy = wmax(ob[4], hy); hz = wmax(ob[5], hz);
        }
        h = 1 + (h > sh ? h : sh);
        lc = L == p ? n - 1 + p : p;
        rc = p + 1 == R ? n + p : p + 1;
        l = L; r = R;
        if (l == 0 && r == n - 1) {
            put_node(&nodes[0], lx, ly, lz, node0 + (uint32_t)lc, hx, hy, hz, node0 + (uint32_t)rc);
            atomicMax(depth_out, h);
            break;
        }
        left = s_d[r - B0 + 1] > s_d[l - B0];
    }
    __syncthreads();
    if (t == 0) esc_count[blockIdx.x] = s_nesc;
}

// BUF: the records through a buffer resource (16-byte device-scope accesses)
template <bool BUF>
__global__ __launch_bounds__(BLOCK) void k_lbvh_top(const uint32_t *__restrict__ codes, const uint32_t *__restrict__ items,
                                                      const float4 *__restrict__ boxes, int n, int groups, uint32_t node0,
                                                      LrNode *nodes, unsigned long long *other, const uint4 *__restrict__ esc,
                                                      const uint32_t *__restrict__ esc_count, uint32_t *depth_out)
{
    const int g = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);   // one wave per group's list
    if (g >= groups) return;
    const uint32_t cnt = esc_count[g];
    const __amdgpu_buffer_rsrc_t rs = node_rsrc(nodes);
    for (uint32_t k = threadIdx.x & 63; k < cnt; k += 64) {
        const uint4 q = esc[(size_t)g * LBVH_SLOTS + k];
        int l = (int)q.x, r = (int)q.y;
        uint32_t h = q.z;
        bool left = is_left(codes, n, l, r);
        const int me = l == r ? n - 1 + l : (left ? r : l);
        const float4 *rec = reinterpret_cast<const float4 *>(&nodes[me]);   // written by k_lbvh
        const float4 blo = rec[0], bhi = rec[1];
        float lx = blo.x, ly = blo.y, lz = blo.z, hx = bhi.x, hy = bhi.y, hz = bhi.z;
        for (;;) {
            const int p = left ? r : l - 1;             // the parent's split
            const int end = left ? l : r;
            // The arrival hands this subtree's record over. The records move as device-scope
            // accesses (sc1 stores and loads, coherent across the XCDs once done), so the
            // exchange needs only this thread's stores done before it (wait_mem: the wait, and
            // compiler barriers on both sides), and the second to arrive reads the first's record
            // at an address taken from the exchanged value, after it. An acquire-release exchange
            // instead writes back and invalidates the L2 at every climb: measured 1.68 -> 6.9 ms
            // for a 16M build.
            wait_mem();
            const unsigned long long mine = ((unsigned long long)h << 32) | ((uint32_t)end + 1u);
            const unsigned long long o = __hip_atomic_exchange(&other[p], mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o == 0) break;                          // the sibling is not done: it goes on
            const int oend = (int)(uint32_t)o - 1;
            const uint32_t sh = (uint32_t)(o >> 32);
            const int L = left ? end : oend, R = left ? oend : end;
            const int lc = L == p ? n - 1 + p : p, rc = p + 1 == R ? n + p : p + 1;
            const int sib = left ? rc : lc;
            float sx, sy, sz, tx, ty, tz;
            if (sib >= n - 1) {   // a leaf: its box is the input's
                const uint32_t it = items[sib - (n - 1)];
                const float4 a = boxes[2 * (size_t)it], b = boxes[2 * (size_t)it + 1];
                sx = a.x; sy = a.y; sz = a.z; tx = b.x; ty = b.y; tz = b.z;
            } else if (BUF) {
                const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, sib * 32, 0, CPOL_SC1);
                const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, sib * 32 + 16, 0, CPOL_SC1);
                sx = __uint_as_float(a.x); sy = __uint_as_float(a.y); sz = __uint_as_float(a.z);
                tx = __uint_as_float(b.x); ty = __uint_as_float(b.y); tz = __uint_as_float(b.z);
            } else {
                const uint64_t *sp = reinterpret_cast<const uint64_t *>(&nodes[sib]);
                const uint64_t q0 = ld_dev64(sp), q1 = ld_dev64(sp + 1), q2 = ld_dev64(sp + 2), q3 = ld_dev64(sp + 3);
                sx = __uint_as_float((uint32_t)q0); sy = __uint_as_float((uint32_t)(q0 >> 32));
                sz = __uint_as_float((uint32_t)q1);
                tx = __uint_as_float((uint32_t)q2); ty = __uint_as_float((uint32_t)(q2 >> 32));
                tz = __uint_as_float((uint32_t)q3);
            }
            float nlx, nly, nlz, nhx, nhy, nhz;
            if (left) {
                nlx = wmin(lx, sx); nly = wmin(ly, sy); nlz = wmin(lz, sz);
                nhx = wmax(hx, tx); nhy = wmax(hy, ty); nhz = wmax(hz, tz);
            } else {
                nlx = wmin(sx, lx); nly = wmin(sy, ly); nlz = wmin(sz, lz);
                nhx = wmax(tx, hx); nhy = wmax(ty, hy); nhz = wmax(tz, hz);
            }
            lx = nlx; ly = nly; lz = nlz; hx = nhx; hy = nhy; hz = nhz;
            h = 1 + (h > sh ? h : sh);
            l = L; r = R;
            if (l == 0 && r == n - 1) {
                put_node(&nodes[0], lx, ly, lz, node0 + (uint32_t)lc, hx, hy, hz, node0 + (uint32_t)rc);
                atomicMax(depth_out, h);
                break;
            }
            left = is_left(codes, n, l, r);
            const int mi = left ? r : l;
            if (BUF) put_node_buf(rs, mi, lx, ly, lz, node0 + (uint32_t)lc, hx, hy, hz, node0 + (uint32_t)rc);
            else put_node_dev(&nodes[mi], lx, ly, lz, node0 + (uint32_t)lc, hx, hy, hz, node0 + (uint32_t)rc);
        }
    }
}

// ---- PLOC ------------------------------------------------------------------------------------
struct PlocState {
    uint32_t nc;          // clusters in the current list
    uint32_t parity;      // which of the two lists is current
    uint64_t live;        // the live clusters summed over the rounds (the build's traffic)
    uint32_t pad[60];
    uint32_t ctr[PLOC_MAX_ROUNDS];   // each round's tile counter
};

// area (trace_ploc.wgsl:18), no fused multiply-add: ((ex*ey) + (ey*ez)) + (ez*ex)
__device__ __forceinline__ float half_area(float lx, float ly, float lz, float hx, float hy, float hz)
{
    const float ex = hx - lx, ey = hy - ly, ez = hz - lz;
    const float a = ex * ey, b = ey * ez, c = ez * ex;
    const float s = a + b;
    return s + c;
}

// nearer (trace_ploc.wgsl:60-66)
__device__ __forceinline__ bool nearer(float d, uint32_t i, uint32_t j, float bd, uint32_t bj)
{
    if (bj == LR_NONE || d < bd) return true;
    if (d > bd) return false;
    const uint32_t mn = min(i, j), bmn = min(i, bj);
    return mn < bmn || (mn == bmn && max(i, j) < max(i, bj));
}

// final_node (trace_ploc.wgsl:33-38) of cluster x with box (lo, hi)
__device__ __forceinline__ void put_final(LrNode *slot, uint32_t x, float4 lo, float4 hi, uint32_t n, uint32_t node0, uint32_t ref0)
{
    if (x < n - 1u) put_node(slot, lo.x, lo.y, lo.z, node0 + 1u + 2u * x, hi.x, hi.y, hi.z, node0 + 2u + 2u * x);
    else put_node(slot, lo.x, lo.y, lo.z, ref0 + x - (n - 1u), hi.x, hi.y, hi.z, 0x80000001u);
}

// merge (trace_ploc.wgsl:39-46): clusters c (the lower position) and d into internal k
__device__ __forceinline__ void ploc_merge(uint32_t k, uint32_t c, float4 clo, float4 chi, uint32_t d, float4 dlo, float4 dhi,
                                           uint32_t n, uint32_t node0, uint32_t ref0, float4 *bx, LrNode *nodes, uint32_t *height)
{
    const float4 lo = make_float4(wmin(clo.x, dlo.x), wmin(clo.y, dlo.y), wmin(clo.z, dlo.z), 0.0f);
    const float4 hi = make_float4(wmax(chi.x, dhi.x), wmax(chi.y, dhi.y), wmax(chi.z, dhi.z), 0.0f);
    bx[2 * (size_t)k] = lo;
    bx[2 * (size_t)k + 1] = hi;
    put_final(&nodes[1 + 2 * (size_t)k], c, clo, chi, n, node0, ref0);
    put_final(&nodes[2 + 2 * (size_t)k], d, dlo, dhi, n, node0, ref0);
    const uint32_t hc = height[c], hd = height[d];
    height[k] = 1 + (hc > hd ? hc : hd);
    if (k == 0) put_node(&nodes[0], lo.x, lo.y, lo.z, node0 + 1u, hi.x, hi.y, hi.z, node0 + 2u);
}

// init (trace_ploc.wgsl:68-82): the leaves in Morton order are the first clusters
__global__ __launch_bounds__(BLOCK) void k_ploc_init(const uint32_t *__restrict__ items, const float4 *__restrict__ boxes, int n,
                                                       uint32_t ref0, uint32_t ref_add, uint32_t *aux, float4 *bx, uint32_t *list0,
                                                       uint32_t *height, PlocState *st, uint64_t *status, size_t status_words)
{
    const size_t g = blockIdx.x * (size_t)BLOCK + threadIdx.x;
    for (size_t i = g; i < status_words; i += (size_t)gridDim.x * BLOCK) status[i] = 0;
    if (g < sizeof(PlocState) / 4) reinterpret_cast<uint32_t *>(st)[g] = g == 0 ? (uint32_t)n : 0u;
    if (g >= (size_t)n) return;
    const int i = (int)g;
    const uint32_t item = items[i];
    const size_t x = (size_t)n - 1 + i;
    bx[2 * x] = boxes[2 * (size_t)item];
    bx[2 * x + 1] = boxes[2 * (size_t)item + 1];
    aux[i] = item + ref_add;
    height[x] = 1;
    list0[i] = (uint32_t)x;
}

// the nearest of position i (< nc) among the window held in LDS: wlo/whi[l] is position base - 2R + l
__device__ __forceinline__ uint32_t nearest_in(int i, int base, uint32_t nc, const float4 *wlo, const float4 *whi)
{
    const int li = i - base + 2 * PLOC_R;
    const float4 lo = wlo[li], hi = whi[li];
    const int j0 = i - PLOC_R > 0 ? i - PLOC_R : 0;
    const int j1 = i + PLOC_R < (int)nc - 1 ? i + PLOC_R : (int)nc - 1;
    float bd = 0.0f;
    uint32_t bj = LR_NONE;
    for (int j = j0; j <= j1; j++) {
        if (j == i) continue;
        const int lj = j - base + 2 * PLOC_R;
        const float4 ol = wlo[lj], oh = whi[lj];
        const float d = half_area(wmin(lo.x, ol.x), wmin(lo.y, ol.y), wmin(lo.z, ol.z),
                                  wmax(hi.x, oh.x), wmax(hi.y, oh.y), wmax(hi.z, oh.z));
        if (nearer(d, (uint32_t)i, (uint32_t)j, bd, bj)) { bd = d; bj = (uint32_t)j; }
    }
    return bj;
}

// One PLOC round (nn + mark + top + scatter of trace_ploc.wgsl:84-234 in one kernel): tiles of
// 256 positions taken in order from a counter; each loads the boxes of its positions +- 2R
// into LDS, finds the nearest of positions +- R, marks who stays and who merges, takes its
// place in the next list from the tiles before it (look-back), writes the merged nodes and
// its part of the next list. Blocks loop over tiles, so any grid size covers the round.
__global__ __launch_bounds__(BLOCK, 7) void k_ploc_round(int round, int n, uint32_t small, uint32_t node0, uint32_t ref0,
                                                        float4 *bx, uint32_t *list0, uint32_t *list1, LrNode *nodes,
                                                        uint32_t *height, PlocState *st, uint64_t *status0, uint64_t *status1)
{
    __shared__ float4 wlo[PLOC_WIN], whi[PLOC_WIN];
    __shared__ uint32_t wid[PLOC_WIN];
    __shared__ int8_t snn[PLOC_NN];   // the nearest as an offset from the position (within +-R; 0: none)
    __shared__ uint32_t s_wsum[BLOCK / 64];
    __shared__ uint32_t s_tile, s_ek, s_em;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_tile = atomicAdd(&st->ctr[round], 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        const uint32_t nc = ld_dev(&st->nc), p = ld_dev(&st->parity);
        const uint32_t ntiles = (nc + PLOC_TILE - 1) / PLOC_TILE;
        if (nc <= small || tile >= ntiles) return;
        const uint32_t *cur = p ? list1 : list0;
        uint32_t *nxt = p ? list0 : list1;
        uint64_t *status = p ? status1 : status0;
        uint64_t *status_next = p ? status0 : status1;
        if (tid == 0) status_next[tile] = 0;       // (the next round's, cleared for it)
        const int base = (int)tile * PLOC_TILE;
        for (int l = tid; l < PLOC_WIN; l += BLOCK) {
            const int g = base - 2 * PLOC_R + l;
            if (g >= 0 && g < (int)nc) {
                const uint32_t x = cur[g];
                wid[l] = x;
                wlo[l] = bx[2 * (size_t)x];
                whi[l] = bx[2 * (size_t)x + 1];
            }
        }
        __syncthreads();
        for (int q = tid; q < PLOC_NN; q += BLOCK) {
            const int i = base - PLOC_R + q;
            const uint32_t bj = (i >= 0 && i < (int)nc) ? nearest_in(i, base, nc, wlo, whi) : LR_NONE;
            snn[q] = bj == LR_NONE ? (int8_t)0 : (int8_t)((int)bj - i);
        }
        __syncthreads();
        // keep_merge (trace_ploc.wgsl:130-145): PLOC_PER adjacent positions a thread, so that the
        // scans of the threads' sums keep the positions' order
        uint32_t j[PLOC_PER], keep[PLOC_PER], merge[PLOC_PER], kc = 0, mc = 0;
#pragma unroll
        for (int q = 0; q < PLOC_PER; q++) {
            const int o = PLOC_PER * tid + q, i = base + o;
            const bool valid = i < (int)nc;
            const int dj = valid ? (int)snn[o + PLOC_R] : 0;
            j[q] = dj != 0 ? (uint32_t)(i + dj) : LR_NONE;
            bool mu = false;
            if (j[q] < nc) mu = (int)j[q] + (int)snn[(int)j[q] - base + PLOC_R] == i;
            keep[q] = valid && !(mu && (uint32_t)i > j[q]);
            merge[q] = valid && mu && (uint32_t)i < j[q];
            kc += keep[q];
            mc += merge[q];
        }
        uint32_t tk, tm;
        const uint32_t xk = block_excl_scan<BLOCK / 64>(kc, &tk, s_wsum);
        const uint32_t xm = block_excl_scan<BLOCK / 64>(mc, &tm, s_wsum);
        if (tid < 64) {
            // the look-back by the first wave, 64 predecessor tiles at once (lane l: tile t - l):
            // the nearest inclusive prefix among them ends it; a needed tile not yet posted means
            // the same window again
            const uint64_t mine = ((uint64_t)tk << 31) | tm;
            uint64_t ek = 0, em = 0;
            if (tile == 0) {
                if (tid == 0) st_dev64(&status[0], (2ull << 62) | mine);
            } else {
                if (tid == 0) st_dev64(&status[tile], (1ull << 62) | mine);
                for (int t = (int)tile - 1;;) {
                    const int q = t - tid;
                    const uint64_t sv = q >= 0 ? ld_dev64(&status[q]) : (2ull << 62);
                    const uint64_t f = sv >> 62;
                    const uint64_t inc = __ballot(f == 2), zero = __ballot(f == 0);
                    const int stop = inc ? __ffsll((unsigned long long)inc) - 1 : 63;
                    const uint64_t need = stop == 63 ? ~0ull : ((2ull << stop) - 1ull);
                    if (zero & need) {
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    uint64_t ak = tid <= stop ? (sv >> 31) & 0x7fffffffull : 0ull;
                    uint64_t am = tid <= stop ? sv & 0x7fffffffull : 0ull;
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        ak += __shfl_xor(ak, o, 64);
                        am += __shfl_xor(am, o, 64);
                    }
                    ek += ak;
                    em += am;
                    if (inc) break;
                    t -= 64;
                }
                if (tid == 0) st_dev64(&status[tile], (2ull << 62) | ((ek + tk) << 31) | (em + tm));
            }
            if (tid == 0) {
                s_ek = (uint32_t)ek;
                s_em = (uint32_t)em;
                if (tile == ntiles - 1) {              // the round's last tile: the next count
                    st_dev(&st->nc, (uint32_t)ek + tk);
                    st_dev(&st->parity, 1u - p);
                    st_dev64(&st->live, ld_dev64(&st->live) + nc);   // (one writer a round)
                }
            }
        }
        __syncthreads();
        uint32_t dk = s_ek + xk, dm = s_em + xm;
#pragma unroll
        for (int q = 0; q < PLOC_PER; q++) {
            if (!keep[q]) continue;
            const int li = PLOC_PER * tid + q + 2 * PLOC_R;
            const uint32_t c = wid[li];
            uint32_t out = c;
            if (merge[q]) {
                const int lj = (int)j[q] - base + 2 * PLOC_R;
                const uint32_t k = ((uint32_t)n - 2u) - ((uint32_t)n - nc + dm);
                dm++;
                ploc_merge(k, c, wlo[li], whi[li], wid[lj], wlo[lj], whi[lj], (uint32_t)n, node0, ref0, bx, nodes, height);
                out = k;
            }
            nxt[dk++] = out;
        }
        __syncthreads();
    }
}

// small (trace_ploc.wgsl:236-301): the rounds left, in one workgroup, from 1024 clusters (the
// context's ploc_small: below it a round's launch costs more than its work; above it the
// nearest search, bound by one compute unit's VALU, is better spread over the chip). While
// more than SMALL_MAX clusters are left (only when the rounds before could not merge enough:
// boxes that tie everywhere merge one pair a round) a round goes through device memory; from
// SMALL_MAX down, the list, its boxes (six arrays of floats) and heights live in LDS, the
// nearest search reads its 2R neighbours there, and only the merged nodes' records go out.
constexpr int SMALL_BLOCK = 1024;
constexpr int SMALL_MAX = 4096;
constexpr int SMALL_PER = SMALL_MAX / SMALL_BLOCK;   // positions a thread holds in an LDS round

// the nearest of position i among the LDS list of nc (positions +- R; the tie rule of nearer)
__device__ __forceinline__ uint32_t nearest_lds(int i, int nc, const float *lx, const float *ly, const float *lz,
                                                const float *hx, const float *hy, const float *hz)
{
    const float ax = lx[i], ay = ly[i], az = lz[i], bx_ = hx[i], by = hy[i], bz = hz[i];
    const int j0 = i - PLOC_R > 0 ? i - PLOC_R : 0;
    const int j1 = i + PLOC_R < nc - 1 ? i + PLOC_R : nc - 1;
    float bd = 0.0f;
    uint32_t bj = LR_NONE;
    for (int j = j0; j <= j1; j++) {
        if (j == i) continue;
        const float d = half_area(wmin(ax, lx[j]), wmin(ay, ly[j]), wmin(az, lz[j]), wmax(bx_, hx[j]), wmax(by, hy[j]),
                                  wmax(bz, hz[j]));
        if (nearer(d, (uint32_t)i, (uint32_t)j, bd, bj)) { bd = d; bj = (uint32_t)j; }
    }
    return bj;
}

__global__ __launch_bounds__(SMALL_BLOCK) void k_ploc_small(int n, uint32_t node0, uint32_t ref0, float4 *bx, uint32_t *list0,
                                                              uint32_t *list1, uint32_t *nnbuf, LrNode *nodes, uint32_t *height,
                                                              PlocState *st, uint32_t *rounds_out, uint32_t *depth_out)
{
    __shared__ uint32_t s_wsum[SMALL_BLOCK / 64];
    __shared__ float s_lx[SMALL_MAX], s_ly[SMALL_MAX], s_lz[SMALL_MAX], s_hx[SMALL_MAX], s_hy[SMALL_MAX], s_hz[SMALL_MAX];
    __shared__ uint32_t s_id[SMALL_MAX], s_nn[SMALL_MAX], s_h[SMALL_MAX];
    const int tid = threadIdx.x;
    uint32_t nc = st->nc, p = st->parity;
    uint32_t rounds = 0;
    uint64_t live = 0;
    // rounds through device memory while the list does not fit in LDS
    while (nc > (uint32_t)SMALL_MAX) {
        live += nc;
        const uint32_t *cur = p ? list1 : list0;
        uint32_t *nxt = p ? list0 : list1;
        for (int i = tid; i < (int)nc; i += SMALL_BLOCK) {
            const uint32_t xi = cur[i];
            const float4 lo = bx[2 * (size_t)xi], hi = bx[2 * (size_t)xi + 1];
            const int j0 = i - PLOC_R > 0 ? i - PLOC_R : 0;
            const int j1 = i + PLOC_R < (int)nc - 1 ? i + PLOC_R : (int)nc - 1;
            float bd = 0.0f;
            uint32_t bj = LR_NONE;
            for (int j = j0; j <= j1; j++) {
                if (j == i) continue;
                const uint32_t xj = cur[j];
                const float4 ol = bx[2 * (size_t)xj], oh = bx[2 * (size_t)xj + 1];
                const float d = half_area(wmin(lo.x, ol.x), wmin(lo.y, ol.y), wmin(lo.z, ol.z),
                                          wmax(hi.x, oh.x), wmax(hi.y, oh.y), wmax(hi.z, oh.z));
                if (nearer(d, (uint32_t)i, (uint32_t)j, bd, bj)) { bd = d; bj = (uint32_t)j; }
            }
            nnbuf[i] = bj;
        }
        __syncthreads();
        const int per = ((int)nc + SMALL_BLOCK - 1) / SMALL_BLOCK;
        const int s0 = min(tid * per, (int)nc), s1 = min(s0 + per, (int)nc);
        uint32_t ck = 0, cm = 0;
        for (int i = s0; i < s1; i++) {
            const uint32_t j = nnbuf[i];
            const bool mu = j < nc && nnbuf[j] == (uint32_t)i;
            ck += !(mu && (uint32_t)i > j);
            cm += mu && (uint32_t)i < j;
        }
        uint32_t tk, tm;
        uint32_t dk = block_excl_scan<SMALL_BLOCK / 64>(ck, &tk, s_wsum);
        uint32_t dm = block_excl_scan<SMALL_BLOCK / 64>(cm, &tm, s_wsum);
        for (int i = s0; i < s1; i++) {
            const uint32_t j = nnbuf[i];
            const bool mu = j < nc && nnbuf[j] == (uint32_t)i;
            if (mu && (uint32_t)i > j) continue;
            const uint32_t c = cur[i];
            uint32_t out = c;
            if (mu && (uint32_t)i < j) {
                const uint32_t d = cur[j];
                const uint32_t k = ((uint32_t)n - 2u) - ((uint32_t)n - nc + dm);
                dm++;
                ploc_merge(k, c, bx[2 * (size_t)c], bx[2 * (size_t)c + 1], d, bx[2 * (size_t)d], bx[2 * (size_t)d + 1],
                           (uint32_t)n, node0, ref0, bx, nodes, height);
                out = k;
            }
            nxt[dk++] = out;
        }
        __syncthreads();
        // every round merges at least the pair nearest of all (it is mutual); a round that
        // merges nothing (NaN boxes) ends the loop instead of spinning
        if (tk >= nc) {
            nc = 0;
            break;
        }
        nc = tk;
        p = 1u - p;
        rounds++;
    }
    // the rest in LDS
    if (nc > 0) {
        const uint32_t *cur = p ? list1 : list0;
        for (int i = tid; i < (int)nc; i += SMALL_BLOCK) {
            const uint32_t x = cur[i];
            const float4 lo = bx[2 * (size_t)x], hi = bx[2 * (size_t)x + 1];
            s_id[i] = x;
            s_lx[i] = lo.x; s_ly[i] = lo.y; s_lz[i] = lo.z;
            s_hx[i] = hi.x; s_hy[i] = hi.y; s_hz[i] = hi.z;
            s_h[i] = height[x];
        }
        __syncthreads();
    }
    while (nc > 1) {
        live += nc;
        for (int i = tid; i < (int)nc; i += SMALL_BLOCK) s_nn[i] = nearest_lds(i, (int)nc, s_lx, s_ly, s_lz, s_hx, s_hy, s_hz);
        __syncthreads();
        // each thread takes a run of positions (the list's order is kept by the scans), holds
        // what it writes in registers, and writes after every read of the round is done
        const int per = ((int)nc + SMALL_BLOCK - 1) / SMALL_BLOCK;
        const int s0 = min(tid * per, (int)nc), s1 = min(s0 + per, (int)nc);
        uint32_t ck = 0, cm = 0;
        for (int i = s0; i < s1; i++) {
            const uint32_t j = s_nn[i];
            const bool mu = j < nc && s_nn[j] == (uint32_t)i;
            ck += !(mu && (uint32_t)i > j);
            cm += mu && (uint32_t)i < j;
        }
        uint32_t tk, tm;
        const uint32_t dk0 = block_excl_scan<SMALL_BLOCK / 64>(ck, &tk, s_wsum);
        uint32_t dm = block_excl_scan<SMALL_BLOCK / 64>(cm, &tm, s_wsum);
        uint32_t o_id[SMALL_PER], o_h[SMALL_PER];
        float o_b[SMALL_PER][6];
        int m = 0;
#pragma unroll
        for (int q = 0; q < SMALL_PER; q++) {
            const int i = s0 + q;
            if (i >= s1) break;
            const uint32_t j = s_nn[i];
            const bool mu = j < nc && s_nn[j] == (uint32_t)i;
            if (mu && (uint32_t)i > j) continue;
            float b0 = s_lx[i], b1 = s_ly[i], b2 = s_lz[i], b3 = s_hx[i], b4 = s_hy[i], b5 = s_hz[i];
            uint32_t id = s_id[i], h = s_h[i];
            if (mu && (uint32_t)i < j) {
                const uint32_t k = ((uint32_t)n - 2u) - ((uint32_t)n - nc + dm);
                dm++;
                const float4 clo = make_float4(b0, b1, b2, 0.0f), chi = make_float4(b3, b4, b5, 0.0f);
                const float4 dlo = make_float4(s_lx[j], s_ly[j], s_lz[j], 0.0f), dhi = make_float4(s_hx[j], s_hy[j], s_hz[j], 0.0f);
                b0 = wmin(clo.x, dlo.x); b1 = wmin(clo.y, dlo.y); b2 = wmin(clo.z, dlo.z);
                b3 = wmax(chi.x, dhi.x); b4 = wmax(chi.y, dhi.y); b5 = wmax(chi.z, dhi.z);
                put_final(&nodes[1 + 2 * (size_t)k], id, clo, chi, (uint32_t)n, node0, ref0);
                put_final(&nodes[2 + 2 * (size_t)k], s_id[j], dlo, dhi, (uint32_t)n, node0, ref0);
                const uint32_t hd = s_h[j];
                h = 1 + (h > hd ? h : hd);
                if (k == 0) put_node(&nodes[0], b0, b1, b2, node0 + 1u, b3, b4, b5, node0 + 2u);
                id = k;
            }
            o_id[m] = id;
            o_h[m] = h;
            o_b[m][0] = b0; o_b[m][1] = b1; o_b[m][2] = b2; o_b[m][3] = b3; o_b[m][4] = b4; o_b[m][5] = b5;
            m++;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < SMALL_PER; q++) {
            if (q >= m) break;
            const uint32_t d = dk0 + (uint32_t)q;
            s_id[d] = o_id[q];
            s_h[d] = o_h[q];
            s_lx[d] = o_b[q][0]; s_ly[d] = o_b[q][1]; s_lz[d] = o_b[q][2];
            s_hx[d] = o_b[q][3]; s_hy[d] = o_b[q][4]; s_hz[d] = o_b[q][5];
        }
        __syncthreads();
        if (tk >= nc) break;   // merged nothing (NaN boxes): ends instead of spinning
        nc = tk;
        rounds++;
    }
    if (tid == 0) {
        *rounds_out = rounds;
        *depth_out = nc == 1 ? s_h[0] : height[0];
        st->live += live;
    }
}

// ---- the scratch layout of one build --------------------------------------------------------------
struct Layout {
    size_t params, bounds_tmp, hist, ticket, bucket, keys0, vals0, keys1, vals1, zero, zero_words, status,
        ctr, depth, other, height, bx, list0, list1, nnbuf, pst, pstatus, pstatus_words, rounds, total;
};

Layout layout(int n)
{
    Layout L;
    const size_t N = n > 0 ? (size_t)n : 1;
    const size_t tiles = (N + SORT_TILE - 1) / SORT_TILE;
    const size_t ptiles = (N + PLOC_TILE - 1) / PLOC_TILE;
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t at = o; o += lr_align(bytes); return at; };
    L.params = take(sizeof(LrBuildParamsDev));
    L.bounds_tmp = take(BOUNDS_BLOCKS * 6 * 4 + 64);
    L.hist = take(HIST_COPIES * SORT_PASSES * 256 * 4);
    L.ticket = take(64);
    L.bucket = take(SORT_PASSES * 256 * 4);
    L.keys0 = take(N * 4);
    L.vals0 = take(N * 4);
    L.keys1 = take(N * 4);
    L.vals1 = take(N * 4);
    // one region the Morton kernel clears: tile counters, depth, look-back states, arrivals
    L.zero = o;
    L.ctr = take(64);
    L.depth = take(64);
    L.status = take(SORT_PASSES * tiles * 256 * 4);
    L.other = take(N * 8);
    L.zero_words = (o - L.zero) / 4;
    // PLOC's; an LBVH build's listed nodes (k_lbvh's slots, 16 bytes each) and their counts
    const size_t groups = (N + LBVH_GROUP - 1) / LBVH_GROUP;
    L.height = take(std::max(8 * N, groups * 4));
    L.bx = take(std::max((2 * N) * 32, groups * LBVH_SLOTS * 16));
    L.list0 = take(N * 4);
    L.list1 = take(N * 4);
    L.nnbuf = take(N * 4);
    L.pst = take(sizeof(PlocState));
    L.pstatus_words = 2 * (ptiles + 1);
    L.pstatus = take(L.pstatus_words * 8);
    L.rounds = take(64);
    L.total = o;
    return L;
}

// ploc_rounds (trace_gpu.c:563): the rounds from n down to `small`, two more to be safe
int ploc_rounds(int n, int small)
{
    int r = 0;
    for (int64_t m = n; m > small; m = m * 5 / 6) r++;
    return r ? r + 2 : 0;
}

}  // namespace

size_t lr_bounds_tmp_bytes() { return BOUNDS_BLOCKS * 6 * 4 + 64; }

int lr_launch_tri_boxes(LrhContext *c, const float *tris, int n, float *boxes)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_tri_boxes, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->stream,
                       reinterpret_cast<const float4 *>(tris), n, reinterpret_cast<float4 *>(boxes));
    LR_HIP(hipGetLastError());
    return 0;
}

// tmp: BOUNDS_BLOCKS * 6 floats then a zeroed ticket word (at +BOUNDS_BLOCKS*24)
int lr_launch_bounds(LrhContext *c, const float *boxes, int n, LrBuildParamsDev *out, void *tmp)
{
    int blocks = (n + BLOCK - 1) / BLOCK;
    if (blocks > BOUNDS_BLOCKS) blocks = BOUNDS_BLOCKS;
    if (blocks < 1) blocks = 1;
    float *partial = (float *)tmp;
    uint32_t *ticket = (uint32_t *)((char *)tmp + BOUNDS_BLOCKS * 24);
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(BLOCK), 0, c->stream, reinterpret_cast<const float4 *>(boxes), n,
                       partial, ticket, out);
    LR_HIP(hipGetLastError());
    return 0;
}

size_t lr_build_scratch_bytes(int n) { return layout(n).total; }

namespace {

// one build with its scratch at S (layout(n).total bytes, the counters before keys0 at zero)
int build_at(LrhContext *c, char *S, const float *boxes, int n, const float *host_bounds, int builder, uint32_t node0,
             uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux, uint32_t *sorted_codes_out,
             uint32_t *sorted_items_out, LrhBuildStats *stats)
{
    const Layout L = layout(n);
    int rc = 0;
    auto at = [&](size_t off) { return (void *)(S + off); };
    LrBuildParamsDev *prm = (LrBuildParamsDev *)at(L.params);
    uint32_t *k0 = (uint32_t *)at(L.keys0), *v0 = (uint32_t *)at(L.vals0);
    uint32_t *k1 = (uint32_t *)at(L.keys1), *v1 = (uint32_t *)at(L.vals1);
    LrNode *nd = reinterpret_cast<LrNode *>(nodes) + node0;
    uint32_t *ax = aux + ref0;
    hipStream_t s = c->stream;
    if (stats) LR_HIP(hipEventRecord(c->ev[0], s));
    if (host_bounds)
        hipLaunchKernelGGL(k_params, dim3(1), dim3(1), 0, s, prm, host_bounds[0], host_bounds[1], host_bounds[2],
                           host_bounds[3], host_bounds[4], host_bounds[5]);
    else if ((rc = lr_launch_bounds(c, boxes, n, prm, at(L.bounds_tmp)))) return rc;
    const int tiles = (n + SORT_TILE - 1) / SORT_TILE;
    hipLaunchKernelGGL(k_morton, dim3(tiles), dim3(SORT_BLOCK), 0, s, reinterpret_cast<const float4 *>(boxes), n, prm, k0, v0,
                       (uint32_t *)at(L.hist), (uint32_t *)at(L.ticket), (uint32_t *)at(L.bucket), (uint32_t *)at(L.zero),
                       L.zero_words);
    LR_HIP(hipGetLastError());
    uint32_t *ctr = (uint32_t *)at(L.ctr);
    for (int pass = 0; pass < SORT_PASSES; pass++) {
        const uint32_t *ki = pass & 1 ? k1 : k0, *vi = pass & 1 ? v1 : v0;
        uint32_t *ko = pass & 1 ? k0 : k1, *vo = pass & 1 ? v0 : v1;
        hipLaunchKernelGGL(k_sort_pass, dim3(tiles), dim3(SORT_BLOCK), 0, s, ki, vi, ko, vo, n, 8 * pass,
                           (const uint32_t *)at(L.bucket) + 256 * pass,
                           (uint32_t *)at(L.status) + (size_t)pass * tiles * 256, ctr + pass);
        LR_HIP(hipGetLastError());
    }
    // sorted: k0 (codes), v0 (items)
    if (sorted_codes_out) LR_HIP(hipMemcpyAsync(sorted_codes_out, k0, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    if (sorted_items_out) LR_HIP(hipMemcpyAsync(sorted_items_out, v0, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    uint32_t *depth = (uint32_t *)at(L.depth);
    uint32_t *height = (uint32_t *)at(L.height);
    int rounds = 0;
    const bool ploc = builder == LRH_PLOC && n > 2;
    if (!ploc) {
        const int groups = (n + LBVH_GROUP - 1) / LBVH_GROUP;
        uint4 *esc = (uint4 *)at(L.bx);
        uint32_t *esc_count = (uint32_t *)at(L.height);
        hipLaunchKernelGGL(k_lbvh, dim3(groups), dim3(LBVH_GROUP), 0, s, k0, v0, reinterpret_cast<const float4 *>(boxes), n, node0,
                           ref0, ref_add, nd, ax, esc, esc_count, depth);
        LR_HIP(hipGetLastError());
        // the climbs above the groups; the buffer path while the records lie within 2 GiB of
        // the array's base
        const dim3 tgrid((unsigned)((groups + BLOCK / 64 - 1) / (BLOCK / 64)));
        if ((2 * (size_t)n - 1) * 32 < 0x7fffffffu)
            hipLaunchKernelGGL(k_lbvh_top<true>, tgrid, dim3(BLOCK), 0, s, k0, v0, reinterpret_cast<const float4 *>(boxes), n,
                               groups, node0, nd, (unsigned long long *)at(L.other), esc, esc_count, depth);
        else
            hipLaunchKernelGGL(k_lbvh_top<false>, tgrid, dim3(BLOCK), 0, s, k0, v0, reinterpret_cast<const float4 *>(boxes), n,
                               groups, node0, nd, (unsigned long long *)at(L.other), esc, esc_count, depth);
        LR_HIP(hipGetLastError());
    } else {
        float4 *bx = (float4 *)at(L.bx);
        uint32_t *l0 = (uint32_t *)at(L.list0), *l1 = (uint32_t *)at(L.list1);
        PlocState *pst = (PlocState *)at(L.pst);
        uint64_t *ps = (uint64_t *)at(L.pstatus);
        const size_t ptiles = L.pstatus_words / 2;
        hipLaunchKernelGGL(k_ploc_init, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, v0,
                           reinterpret_cast<const float4 *>(boxes), n, ref0, ref_add, ax, bx, l0, height, pst, ps,
                           L.pstatus_words);
        LR_HIP(hipGetLastError());
        const int small = std::min(c->ploc_small > 0 ? c->ploc_small : 1024, SMALL_MAX);
        int nr = ploc_rounds(n, small);
        if (nr > PLOC_MAX_ROUNDS) nr = PLOC_MAX_ROUNDS;
        double est = n;
        for (int r = 0; r < nr; r++) {
            int grid = (int)((est + PLOC_TILE - 1) / PLOC_TILE);
            if (grid < 16) grid = 16;
            if (grid > 4096) grid = 4096;
            hipLaunchKernelGGL(k_ploc_round, dim3(grid), dim3(BLOCK), 0, s, r, n, (uint32_t)small, node0, ref0, bx, l0, l1,
                               nd, height, pst, ps, ps + ptiles);
            LR_HIP(hipGetLastError());
            est *= 0.84;
        }
        hipLaunchKernelGGL(k_ploc_small, dim3(1), dim3(SMALL_BLOCK), 0, s, n, node0, ref0, bx, l0, l1,
                           (uint32_t *)at(L.nnbuf), nd, height, pst, (uint32_t *)at(L.rounds), depth);
        LR_HIP(hipGetLastError());
        rounds = nr;
    }
    if (stats) {
        LR_HIP(hipEventRecord(c->ev[1], s));
        LR_HIP(hipEventSynchronize(c->ev[1]));
        float ms = 0;
        LR_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        uint32_t dep = 0, small_rounds = 0;
        uint64_t live = 0;
        LR_HIP(hipMemcpy(&dep, depth, 4, hipMemcpyDeviceToHost));
        if (ploc) {
            LR_HIP(hipMemcpy(&small_rounds, at(L.rounds), 4, hipMemcpyDeviceToHost));
            LR_HIP(hipMemcpy(&live, (const char *)at(L.pst) + offsetof(PlocState, live), 8, hipMemcpyDeviceToHost));
        }
        stats->clusters = live;
        stats->refitted = 0;
        stats->depth = (int)dep;
        stats->rounds = ploc ? rounds + (int)small_rounds : 0;
        stats->ms = ms;
        c->last_build_ms = ms;
    }
    return 0;
}

}  // namespace

int lr_build_run(LrhContext *c, const float *boxes, int n, const float *host_bounds, int builder, uint32_t node0,
                 uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux, uint32_t *sorted_codes_out,
                 uint32_t *sorted_items_out, LrhBuildStats *stats)
{
    int rc = lr_scratch_reserve(c, LR_SCRATCH_BUILD, layout(n).total);
    if (rc) return rc;
    return build_at(c, (char *)c->scratch[LR_SCRATCH_BUILD].base, boxes, n, host_bounds, builder, node0, ref0, ref_add,
                    nodes, aux, sorted_codes_out, sorted_items_out, stats);
}

// ---- the build without allocation (lrh_build_noalloc) --------------------------------------------
// Its scratch is this file-scope array: each device holds its own copy from the library's load
// until the process ends. In host code `scratch` names a host-side stand-in, not the device
// copy: the device address is taken once per device (lr_noalloc_init, from lrh_create, with
// that device current) and only that address is handed to kernels.
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr size_t NOALLOC_BYTES = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static char *g_noalloc_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_noalloc_mu;

namespace {

// clears n words (the counters that must start at zero: a kernel, so a caller may capture it)
__global__ __launch_bounds__(256) void k_clear_words(uint32_t *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0u;
}

}  // namespace

int lr_noalloc_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return lr_fail(LRH_EINVAL, "device %d: the no-allocation build keeps %d devices", device, NOALLOC_MAX_DEVICES);
    if (layout(LRH_NOALLOC_MAX_ITEMS).total > NOALLOC_BYTES)
        return lr_fail(LRH_EDEVICE, "the no-allocation build's scratch (%zu bytes) is below its %d items' %zu",
                       NOALLOC_BYTES, LRH_NOALLOC_MAX_ITEMS, layout(LRH_NOALLOC_MAX_ITEMS).total);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return lr_fail(LRH_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    }
    g_noalloc_ptr[device] = (char *)ptr;
    return 0;
}

int lr_build_noalloc(LrhContext *c, const float *boxes, int n, const float *host_bounds, int builder, uint32_t node0,
                     uint32_t ref0, uint32_t ref_add, uint32_t *nodes, uint32_t *aux, uint32_t *sorted,
                     LrhBuildStats *stats)
{
    char *S = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (c->device >= 0 && c->device < NOALLOC_MAX_DEVICES) S = g_noalloc_ptr[c->device];
    }
    if (!S) return lr_fail(LRH_EDEVICE, "lrh_build_noalloc: device %d has no scratch address", c->device);
    const Layout L = layout(n);
    if (L.total > NOALLOC_BYTES)
        return lr_fail(LRH_EREFUSED, "lrh_build_noalloc: %d items need %zu bytes of scratch, above its %zu", n, L.total,
                       NOALLOC_BYTES);
    // the counters before the keys start at zero (a call that ends normally leaves them so; this
    // clear makes the call independent of any earlier one)
    hipLaunchKernelGGL(k_clear_words, dim3(16), dim3(256), 0, c->stream, (uint32_t *)S, L.keys0 / 4);
    LR_HIP(hipGetLastError());
    return build_at(c, S, boxes, n, host_bounds, builder, node0, ref0, ref_add, nodes, aux, sorted,
                    sorted ? sorted + n : nullptr, stats);
}
