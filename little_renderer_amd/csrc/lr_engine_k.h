// lr_engine_k.h - the engine's four wavefront kernels, templated on the work-group size (and
// the walks on their LDS stack depth). Included once per work-group size by
// lr_engine_b64.hip, lr_engine_b128.hip and lr_engine_b256.hip, each filling the launch table
// (lr_engine.h) for its size.
//
// Light transport: trace_path_sampled (engine/rt/tracer.c:364-509) and trace_image
// (tracer.c:329-356) in the wavefront split of trace_wave.wgsl (wf_gen :51-70, wf_ext :72-84,
// wf_shade :191-310, wf_shadow :335-355, wf_resolve :357-378).
#pragma once

#include "lr_engine.h"
#include "lr_pt_dev.h"

namespace {

// The entries [0, n) in tiles of BLK * mult, claimed by the work-groups one at a time through
// a counter (persistent work-groups loop; a grid that covers n takes one tile each). body(i,
// live) runs for every lane, live or not, so that it may use wave-wide ballots.
template <int BLK, class F>
__device__ __forceinline__ void for_tiles(unsigned int *ctr, uint32_t n, int mult, F &&body)
{
    __shared__ uint32_t s_tile;
    const uint32_t width = (uint32_t)BLK * (uint32_t)mult;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_tile = atomicAdd(ctr, 1u);
        __syncthreads();
        const uint64_t base = (uint64_t)s_tile * width;
        if (base >= n) break;
        for (int j = 0; j < mult; j++) {
            const uint64_t i = base + (uint64_t)j * BLK + threadIdx.x;
            body((uint32_t)i, i < n);
        }
    }
}

// Each kernel clears the other kernels' tile counters as it starts (kernels of one stream run
// one after another, and no kernel follows itself), so every kernel finds its own at zero.
__device__ __forceinline__ void reset_tiles(EngCtl *ctl, int self)
{
    for (int k = 0; k < 4; k++)
        if (k != self) ctl->tiles[k] = 0;
}

__device__ __forceinline__ long long to_fixed(float x)
{
    x = x > LR_FIX_CAP ? LR_FIX_CAP : (x < -LR_FIX_CAP ? -LR_FIX_CAP : x);
    return __float2ll_rn(x * (float)(1u << LR_FIX_SHIFT));
}

// Positions, in entry order, of a tile's outputs to two queues: a scan across the work-group
// and a decoupled look-back over the tiles before it. A tile's status word: 2-bit flag (1: its
// own counts, 2: the counts of all tiles up to it), then two 31-bit counts. Tiles are claimed
// in order by running work-groups, so every tile waited on is being worked on. The last tile
// writes the queues' lengths.
template <int BLK>
__device__ __forceinline__ void ordered_append2(unsigned long long *stat, uint32_t t, uint32_t ntiles, bool fa, bool fb,
                                                uint32_t &pa, uint32_t &pb, unsigned int *len_a, unsigned int *len_b)
{
    constexpr int NW = BLK / 64;
    __shared__ uint32_t s_w[2][NW];
    __shared__ uint32_t s_base[2];
    const uint64_t ma = __ballot(fa), mb = __ballot(fb);
    const int w = (int)(threadIdx.x / 64);
    if (lane_id() == 0) {
        s_w[0][w] = (uint32_t)__popcll(ma);
        s_w[1][w] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    uint32_t pre_a = 0, pre_b = 0, tot_a = 0, tot_b = 0;
    for (int k = 0; k < NW; k++) {
        if (k < w) {
            pre_a += s_w[0][k];
            pre_b += s_w[1][k];
        }
        tot_a += s_w[0][k];
        tot_b += s_w[1][k];
    }
    if (threadIdx.x == 0) {
        unsigned long long ea = 0, eb = 0;
        if (t > 0) {
            st_dev64((uint64_t *)&stat[t], (1ull << 62) | ((unsigned long long)tot_a << 31) | tot_b);
            for (long j = (long)t - 1; j >= 0;) {
                const unsigned long long v = ld_dev64((const uint64_t *)&stat[j]);
                const unsigned f = (unsigned)(v >> 62);
                if (f == 0) continue;
                ea += (v >> 31) & 0x7fffffffull;
                eb += v & 0x7fffffffull;
                if (f == 2) break;
                j--;
            }
        }
        st_dev64((uint64_t *)&stat[t], (2ull << 62) | ((ea + tot_a) << 31) | (eb + tot_b));
        s_base[0] = (uint32_t)ea;
        s_base[1] = (uint32_t)eb;
        if (t == ntiles - 1) {
            *len_a = (unsigned int)(ea + tot_a);
            *len_b = (unsigned int)(eb + tot_b);
        }
    }
    __syncthreads();
    pa = s_base[0] + pre_a + (uint32_t)__popcll(ma & lanemask_lt());
    pb = s_base[1] + pre_b + (uint32_t)__popcll(mb & lanemask_lt());
}

// A path ends: its sample joins its pixel (NaN samples skipped and the clamp applied as
// trace_image does, tracer.c:347-349), its place in the pool is freed, its job counts it. Every
// lane of the wave calls this (on: the lane's path ends).
__device__ __forceinline__ void retire(const EngArgs &A, bool on, uint32_t jslot, uint32_t q, float4 r, bool glow = false)
{
    int pk = -1;          // a pilot job's pass of this sample (its first 4), else -1
    long long pl = 0;     // ... and the sample's luminance, fixed point
    if (on) {
        const JobDev &J = A.jobs[jslot];
        V3 c = v3(r.x, r.y, r.z);
        const bool sane = c.x == c.x && c.y == c.y && c.z == c.z;
        if (J.pilot && q / (uint32_t)J.npix < LR_PILOT_PASSES) pk = (int)(q / (uint32_t)J.npix);
        if (sane) {
            const float m = vmax3(c);
            if (J.clamp > 0.0f && m > J.clamp) c = vscale(c, J.clamp / m);
            const uint32_t pix = q % (uint32_t)J.npix;
            long long *px = J.accum + 3 * (size_t)pix;
            const long long fx = to_fixed(c.x), fy = to_fixed(c.y), fz = to_fixed(c.z);
            if (fx) atomicAdd((unsigned long long *)&px[0], (unsigned long long)fx);
            if (fy) atomicAdd((unsigned long long *)&px[1], (unsigned long long)fy);
            if (fz) atomicAdd((unsigned long long *)&px[2], (unsigned long long)fz);
            if (pk >= 0) pl = to_fixed(lum(c));
            const float4 ga = J.moments ? J.guide[2 * (size_t)pix + 1] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            if (J.moments && !(glow && ga.w == 0.0f)) {
                // the denoiser's moments: the luminance of the sample over the guide's albedo on a
                // surface (its irradiance), and its square (trace.wgsl:1301-1309); not a sample whose
                // camera ray found a bright glow where the guide sees a surface (its pixel's edge
                // with a lamp: it would make the whole pixel look noisy)
                const V3 ir = ga.w == 0.0f ? v3(c.x / rfmax(ga.x, 1e-3f), c.y / rfmax(ga.y, 1e-3f), c.z / rfmax(ga.z, 1e-3f)) : c;
                const float l = lum(ir);
                const float l2 = rfmin(l * l, LR_MOM2_CAP);
                long long *mo = J.moments + 2 * (size_t)pix;
                const long long fl = to_fixed(l), fl2 = __float2ll_rn(l2 * (float)(1u << LR_MOM2_SHIFT));
                if (fl) atomicAdd((unsigned long long *)&mo[0], (unsigned long long)fl);
                if (fl2) atomicAdd((unsigned long long *)&mo[1], (unsigned long long)fl2);
            }
        }
    }
    const uint64_t m = __ballot(on);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    if ((int)lane_id() == leader) atomicAdd(&A.ctl->tail, (unsigned long long)__popcll(m));
    // a pilot job's first passes: each pass's luminance summed over the wave, one atomic per
    // (job, pass) present (added before the snapshot after the graph: complete when read)
    uint64_t pp = __ballot(pk >= 0);
    while (pp) {
        const int l = __ffsll((unsigned long long)pp) - 1;
        const uint32_t j = __shfl(jslot, l, 64);
        const int k = __shfl(pk, l, 64);
        const bool mine = pk == k && jslot == j;
        pp &= ~__ballot(mine);
        long long v = mine ? pl : 0;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((int)lane_id() == l) atomicAdd((unsigned long long *)&A.jpass[j][k], (unsigned long long)v);
    }
    // the job's count of retired paths: one atomic per job present in the wave
    uint64_t pend = m;
    while (pend) {
        const int l = __ffsll((unsigned long long)pend) - 1;
        const uint32_t j = __shfl(jslot, l, 64);
        const uint64_t same = __ballot(on && jslot == j);
        pend &= ~same;
        if ((int)lane_id() == l) {
            const unsigned long long k = (unsigned long long)__popcll(same);
            const unsigned long long old = atomicAdd(&A.jret[j], k);
            if (old + k == A.jobs[j].total) A.jdone[j] = wall_clock64();
        }
    }
}

// ---- two-level scenes (trace.wgsl:186-392: instances, to_object, unit sphere and box) ----
__device__ __forceinline__ void to_object(const float4 *rows, V3 o, V3 d, V3 &oo, V3 &od)
{
    const float4 a = rows[0], b = rows[1], c = rows[2];
    oo = v3(a.x * o.x + a.y * o.y + a.z * o.z + a.w, b.x * o.x + b.y * o.y + b.z * o.z + b.w,
            c.x * o.x + c.y * o.y + c.z * o.z + c.w);
    od = v3(a.x * d.x + a.y * d.y + a.z * d.z, b.x * d.x + b.y * d.y + b.z * d.z, c.x * d.x + c.y * d.y + c.z * d.z);
}
// an object-space normal in the world: by the inverse transpose, normalised
__device__ __forceinline__ V3 normal_to_world(const float4 *rows, V3 n)
{
    const float4 a = rows[0], b = rows[1], c = rows[2];
    return vnorm(v3(a.x * n.x + b.x * n.y + c.x * n.z, a.y * n.x + b.y * n.y + c.y * n.z,
                    a.z * n.x + b.z * n.y + c.z * n.z));
}
__device__ __forceinline__ float hit_unit_sphere(V3 o, V3 d, float tmin, float tmax)
{
    const float a = vdot(d, d), hh = -vdot(d, o), c = vdot(o, o) - 1.0f;
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
__device__ __forceinline__ float hit_unit_box(V3 o, V3 d, float tmin, float tmax)
{
    const V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    const float t0x = (-1.0f - o.x) * inv.x, t0y = (-1.0f - o.y) * inv.y, t0z = (-1.0f - o.z) * inv.z;
    const float t1x = (1.0f - o.x) * inv.x, t1y = (1.0f - o.y) * inv.y, t1z = (1.0f - o.z) * inv.z;
    const float ax = t0x < t1x ? t0x : t1x, ay = t0y < t1y ? t0y : t1y, az = t0z < t1z ? t0z : t1z;
    const float bx = t0x > t1x ? t0x : t1x, by = t0y > t1y ? t0y : t1y, bz = t0z > t1z ? t0z : t1z;
    float enter = ax > ay ? ax : ay;
    enter = enter > az ? enter : az;
    float leave = bx < by ? bx : by;
    leave = leave < bz ? leave : bz;
    if (leave < enter) return -1.0f;
    if (enter > tmin && enter < tmax) return enter;
    if (leave > tmin && leave < tmax) return leave;
    return -1.0f;
}
__device__ __forceinline__ V3 unit_box_normal(V3 p)
{
    const float ax = fabsf(p.x), ay = fabsf(p.y), az = fabsf(p.z);
    if (ax >= ay && ax >= az) return v3(p.x < 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f);
    if (ay >= az) return v3(0.0f, p.y < 0.0f ? -1.0f : 1.0f, 0.0f);
    return v3(0.0f, 0.0f, p.z < 0.0f ? -1.0f : 1.0f);
}

// The closest (ANY: the first) hit in a two-level scene, on the 4-wide trees (walk4's compressed
// nodes, one entry at a time): the top level's wide nodes, whose leaf references hold
// instances, and each mesh's. A mesh instance is entered (the ray taken to its space, EXIT
// pushed under its mesh's root); the hit's b1 carries the instance (as u32). A node's children hit are
// sorted by their boxes' entry distance, whatever they are (a node, a mesh's leaf, an instance
// to enter), so all are visited near to far; unit spheres and boxes are tested on the spot. A
// mesh's triangles are read in leaf order (ltris) beside their aux entries.
template <bool ANY, bool COUNT, int STK>
__device__ WalkHit walk4_tl(const JobDev &J, V3 o, V3 d, float tmin, float tmax, const WalkStack &S, uint32_t *overflow,
                            WalkCnt &c)
{
    // the ray in the space being walked (the world's, or an instance's); the world's inverse
    // direction is computed again on leaving an instance rather than held
    V3 ro = o, rd = d;
    V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    V3 oi = v3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
    uint32_t cur = LR_NONE;
    WalkHit best{tmax, LR_NONE, __uint_as_float(LR_NONE), 0.0f};
    int sp = 0;
    auto push = [&](uint32_t x) {
        if (sp < STK) S.stk[sp * S.stride] = x;
        else S.spill[(size_t)(sp - STK) * S.sstride] = x;
        sp++;
    };
    auto pop = [&](uint32_t &x) -> bool {
        if (sp == 0) return false;
        sp--;
        x = stk_get<STK>(S, sp);
        return true;
    };
    // the entry being visited: a top-level node k, INST | i, BLAS | mesh node, MLEAF | (count -
    // 1) << 24 | first maux entry, EXIT
    constexpr uint32_t MLEAF = 0xc0000000u;
    uint32_t e = 0u;
    for (;;) {
        if (e == LR_WALK_EXIT) {
            cur = LR_NONE;
            ro = o;
            rd = d;
            inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
            oi = v3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
            if (!pop(e)) break;
            continue;
        }
        const uint32_t tag = e & 0xc0000000u;
        if (tag == LR_WALK_INST) {
            const uint32_t i = e & 0x3fffffffu;
            const uint32_t kind = __float_as_uint(J.inst[4 * (size_t)i + 3].x);
            if (kind >= LR_UNIT_SPHERE) {   // a unit sphere or box: tested in its space, the ray kept
                if (COUNT) count_shape(c);
                V3 oo, od;
                to_object(J.inst + 4 * (size_t)i, o, d, oo, od);
                const float th = kind == LR_UNIT_SPHERE ? hit_unit_sphere(oo, od, tmin, best.t) : hit_unit_box(oo, od, tmin, best.t);
                if (th > 0.0f) {
                    best = WalkHit{th, kind, __uint_as_float(i), 0.0f};
                    if (ANY) return best;
                }
                if (!pop(e)) break;
                continue;
            }
            to_object(J.inst + 4 * (size_t)i, o, d, ro, rd);
            inv = v3(safe_inv(rd.x), safe_inv(rd.y), safe_inv(rd.z));
            oi = v3(ro.x * inv.x, ro.y * inv.y, ro.z * inv.z);
            cur = i;
            if (sp + 1 > S.cap) {   // cannot happen with a correct depth: said, never silently dropped
                atomicOr(overflow, 1u);
                break;
            }
            push(LR_WALK_EXIT);
            e = LR_WALK_BLAS | __float_as_uint(J.inst[4 * (size_t)i + 3].w);
            continue;
        }
        if (tag == MLEAF) {   // a mesh's leaf, in the instance's space
            const uint32_t cnt = ((e >> 24) & 0x3fu) + 1u, first = e & 0xffffffu;
            for (uint32_t j = 0; j < cnt; j++) {
                if (COUNT) count_shape(c);
                const uint32_t x = first + j;
                const uint32_t sid = J.maux[x];
                const float4 ta = J.ltris[3 * (size_t)x], tb = J.ltris[3 * (size_t)x + 1], tc = J.ltris[3 * (size_t)x + 2];
                float th, u, v;
                if (hit_tri(ta, tb, tc, ro, rd, tmin, best.t, th, u, v)) {
                    best = WalkHit{th, sid, __uint_as_float(cur), 0.0f};
                    if (ANY) return best;
                }
            }
            if (!pop(e)) break;
            continue;
        }
        const bool blas = tag == LR_WALK_BLAS;
        if (COUNT) count_node(c);
        float t[4];
        uint32_t ref[4];
        wide_test(blas ? J.mnodes4 : J.nodes, e & 0x3fffffffu, oi, inv, tmin, best.t, t, ref);
        float s[4] = {3e38f, 3e38f, 3e38f, 3e38f};
        uint32_t c[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!(t[k] < 3e38f)) continue;
            if (!(ref[k] & LR_LEAF)) {
                c[k] = (blas ? LR_WALK_BLAS : 0u) | ref[k];
                s[k] = t[k];
                continue;
            }
            if (blas) {
                c[k] = MLEAF | (ref[k] & 0x3fffffffu);
                s[k] = t[k];
                continue;
            }
            // a top-level leaf: its instances, entries (one leaf, one instance: sorted with the rest)
            const uint32_t cnt = ((ref[k] >> 24) & 0x3fu) + 1u, first = ref[k] & 0xffffffu;
            if (cnt == 1) {
                c[k] = LR_WALK_INST | J.aux[first];
                s[k] = t[k];
                continue;
            }
            if (sp + cnt > (uint32_t)S.cap) {   // cannot happen with a correct depth: said, never silently dropped
                atomicOr(overflow, 1u);
                return best;
            }
            for (uint32_t j = 0; j < cnt; j++) push(LR_WALK_INST | J.aux[first + j]);
        }
        // nearest first
        float s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3];
        uint32_t c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
        cswap_(s0, c0, s1, c1);
        cswap_(s2, c2, s3, c3);
        cswap_(s0, c0, s2, c2);
        cswap_(s1, c1, s3, c3);
        cswap_(s1, c1, s2, c2);
        const int m = (s0 < 3e38f) + (s1 < 3e38f) + (s2 < 3e38f) + (s3 < 3e38f);
        if (m == 0) {
            if (!pop(e)) break;
            continue;
        }
        if (sp + m - 1 > S.cap) {   // cannot happen with a correct depth: said, never silently dropped
            atomicOr(overflow, 1u);
            break;
        }
        if (m > 3) push(c3);
        if (m > 2) push(c2);
        if (m > 1) push(c1);
        e = c0;
    }
    return best;
}

// regen (wf_gen, trace_wave.wgsl:51-70, with trace_image's seeds, tracer.c:335-346): `claimed`
// free places of the pool start paths next .. next + claimed - 1, each a camera ray and its state
// at the end of the ray queue
template <int BLK>
__global__ __launch_bounds__(BLK) void k_regen(EngArgs A)
{
    EngCtl *ctl = A.ctl;
    const unsigned long long head = ctl->head, tail = ctl->tail, next = ctl->next, iend = ctl->issue_end;
    const unsigned long long lo = ctl->iq_lo, hi = ctl->iq_hi;
    const unsigned long long free_n = tail - head, todo = iend - next;
    const uint32_t n = (uint32_t)(free_n < todo ? free_n : todo);
    const int cur = (int)(ctl->iter & 1ull);
    const uint32_t start = ctl->count[cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->claimed = n;
        reset_tiles(ctl, K_REGEN);
    }
    for_tiles<BLK>(&ctl->tiles[K_REGEN], n, A.mult, [&](uint32_t i, bool live) {
        if (live) {
            const unsigned long long g = next + i;
            // the issue record holding g: the last one that starts at or before it
            unsigned long long a = lo, b = hi;   // records [a, b); the answer in [a, b)
            while (b - a > 1) {
                const unsigned long long mid = a + (b - a) / 2;
                if (A.iq_start[mid & (LR_ISSUE - 1)] <= g) a = mid;
                else b = mid;
            }
            const uint32_t jslot = A.iq_job[a & (LR_ISSUE - 1)] & (LR_JOBS - 1);
            const JobDev &J = A.jobs[jslot];
            const uint32_t q = A.iq_q0[a & (LR_ISSUE - 1)] + (uint32_t)(g - A.iq_start[a & (LR_ISSUE - 1)]);
            const uint32_t pixel = q % (uint32_t)J.npix;
            const uint32_t k = J.s0 + q / (uint32_t)J.npix;
            const int x = (int)(pixel % (uint32_t)J.W), y = (int)(pixel / (uint32_t)J.W);
            const uint64_t seed = ((uint64_t)J.seed_hi << 32) | J.seed_lo;
            uint64_t st = pcg_seed((uint64_t)pixel, k, seed);
            const uint64_t inc = ((uint64_t)pixel << 1u) | 1u;
            const Smp sm{x, y, k, (uint32_t)seed * 0x9e3779b9u, A.mask, &st, inc};
            float jx, jy;
            s2(sm, 0, jx, jy);
            const float u = ((float)x + jx) / (float)J.W, v = ((float)y + jy) / (float)J.H;
            // camera_ray (camera.h:38-42)
            const V3 org = v3(J.cam[0], J.cam[1], J.cam[2]);
            const V3 ul = v3(J.cam[3], J.cam[4], J.cam[5]), hz = v3(J.cam[6], J.cam[7], J.cam[8]),
                     vt = v3(J.cam[9], J.cam[10], J.cam[11]);
            const V3 target = vsub(vadd(ul, vscale(hz, u)), vscale(vt, v));
            const V3 d = vnorm(vsub(target, org));
            // every started path has its camera ray and its state: in claim order, after the
            // queue's paths
            const uint32_t at = start + i;
            A.qray[cur][2 * (size_t)at] = make_float4(org.x, org.y, org.z, J.eps);
            A.qray[cur][2 * (size_t)at + 1] = make_float4(d.x, d.y, d.z, 1e30f);
            A.qid[cur][at] = jslot << LR_SLOT_BITS;
            A.thr[cur][at] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            A.rad[cur][at] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(EMIT_FULL));
            A.rng[cur][at] = st;
            A.meta[cur][at] = make_uint2(q, 0u);
        }
    });
    // the last work-group to finish sets the queue's length (all have read the old one)
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&ctl->regen_done, 1u) == gridDim.x - 1) {
        ctl->count[cur] = start + n;
        ctl->regen_done = 0;
    }
}

// a counting walk's counters summed over the kernel: a wave's lanes' sums, then one add each
// (c: node records, shapes, the wave's node steps, its shape steps; u: its uniform node steps)
__device__ __forceinline__ void add_counts(unsigned long long *c, unsigned long long *u, const WalkCnt &w)
{
    unsigned long long v[5] = {w.n, w.t, w.wn, w.wt, w.wu};
#pragma unroll
    for (int k = 0; k < 5; k++) {
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
        if (lane_id() == 0 && v[k]) atomicAdd(k < 4 ? &c[k] : u, v[k]);
    }
}

// extend (wf_ext, trace_wave.wgsl:72-84): the closest hit of every queued ray, each through
// its own job's tree. Also applies regen's claim to the ring and the issue counter, and clears
// the queues the shade kernel appends to. TL: two-level scenes may be queued (without it, such
// a ray is not walked: flagged, so that the host reports it).
// Without TL the kernel is held to 64 VGPRs, 8 waves per SIMD: a few registers spill (12 bytes a
// lane), and the walk, bound by the latency of its node fetches, gains more from the waves (the
// whole persistent grid of 8 work-groups per compute unit is then resident).
// waves per SIMD a walk kernel is held to: 8, or fewer where its LDS stack (160 KB per compute
// unit) allows fewer anyway
constexpr int walk_waves(int blk, int stk)
{
    return (163840 / (stk * blk * 4 + 4)) * (blk / 64) / 4 < 8 ? (163840 / (stk * blk * 4 + 4)) * (blk / 64) / 4 : 8;
}

template <int BLK, int STK, bool COUNT, bool TL>
__global__ __launch_bounds__(BLK, TL ? 1 : walk_waves(BLK, STK)) void k_extend(EngArgs A)
{
    __shared__ uint32_t stk[STK][BLK];
    EngCtl *ctl = A.ctl;
    const int cur = (int)(ctl->iter & 1ull);
    const uint32_t n = ctl->count[cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned long long c = ctl->claimed;
        ctl->head += c;
        ctl->next += c;
        ctl->claimed = 0;
        ctl->rays_ext += n;
        ctl->count[1 - cur] = 0;
        ctl->count[2] = 0;
        reset_tiles(ctl, K_EXTEND);
    }
    const size_t gtid = (size_t)blockIdx.x * BLK + threadIdx.x;
    // the shade kernel's look-back status words, one per tile of at least 64 entries
    for (size_t k = gtid; k < (size_t)n / 64 + 2; k += (size_t)gridDim.x * BLK) A.ostat[k] = 0ull;
    const WalkStack S{&stk[0][threadIdx.x], BLK, A.spill + gtid, (size_t)A.spill_stride, STK + A.spill_per};
    WalkCnt wc;
    for_tiles<BLK>(&ctl->tiles[K_EXTEND], n, A.mult, [&](uint32_t i, bool live) {
        if (!live) return;
        const uint32_t id = A.qid[cur][i];
        const JobDev &J = A.jobs[id >> LR_SLOT_BITS];
        const float4 r0 = A.qray[cur][2 * (size_t)i], r1 = A.qray[cur][2 * (size_t)i + 1];
        WalkHit h;
        if (TL && J.two_level) h = walk4_tl<false, COUNT, STK>(J, v3of(r0), v3of(r1), r0.w, r1.w, S, A.overflow, wc);
        else if (!J.two_level)
            h = walk4<false, COUNT, STK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, v3of(r0), v3of(r1), r0.w, r1.w, S,
                                         A.overflow, wc);
        else {
            atomicOr(A.overflow, 2u);
            h = WalkHit{r1.w, LR_NONE, 0.0f, 0.0f};
        }
        A.hits[i] = make_float4(h.t, __uint_as_float(h.id), h.b1, h.b2);
    });
    if (COUNT) add_counts(A.counts, A.counts + 8, wc);
}

// ---- walks with lanes refilled (dynamic fetch, Aila & Laine 2009) ----------------------------
// A wave walks its lanes' rays one round at a time (the leaves a lane holds, then one node step:
// walk4's loop body), and once at least A.refill of its 64 lanes have finished, those take the
// next entries of the queue: a wave does not run on for its slowest ray with most of its lanes
// idle. A wave claims the queue 64 entries at a time (the kernel's counter then counts entries,
// not tiles) and hands them to its idle lanes in order. The walk is walk4's (the same visits in
// the same order, the same hit); single-level scenes only. Every lane of a wave reaches the top
// of the round loop together, where the hand-out's ballots and the claim's broadcast are made.

// a lane's walk between rounds
struct WalkState {
    V3 o, d, inv;   // (o * inv, the slab test's other term, made again each node step: fewer registers held)
    float tmin;
    WalkHit best;
    int sp;
    uint32_t cur;
    const float4 *nodes, *ltris, *prims;
    const uint32_t *aux;
    uint32_t nprims;
};

__device__ __forceinline__ void walk_start(WalkState &w, const JobDev &J, float4 r0, float4 r1)
{
    w.o = v3of(r0);
    w.d = v3of(r1);
    w.inv = v3(safe_inv(w.d.x), safe_inv(w.d.y), safe_inv(w.d.z));
    w.tmin = r0.w;
    w.best = WalkHit{r1.w, LR_NONE, 0.0f, 0.0f};
    w.sp = 0;
    w.cur = 0u;
    w.nodes = J.nodes;
    w.aux = J.aux;
    w.ltris = J.ltris;
    w.prims = J.prims;
    w.nprims = J.nprims;
}

// one round of walk4 (ANY: the first hit ends it); true when the walk has ended (w.best its hit)
template <bool ANY, bool COUNT, int STK>
__device__ __forceinline__ bool walk_round(WalkState &w, const WalkStack &S, uint32_t *overflow, WalkCnt &wc)
{
    while (w.cur & LR_LEAF) {
        bool done = false;
        leaf_hits_(w.aux, w.ltris, w.prims, w.nprims, w.cur & 0xffffffu, ((w.cur >> 24) & 0x3fu) + 1u, w.o, w.d, w.tmin, w.best,
                   ANY, done, wc, COUNT);
        if ((ANY && done) || w.sp == 0) return true;
        w.sp--;
        w.cur = stk_get<STK>(S, w.sp);
    }
    if (COUNT) {
        count_node(wc);
        count_uniform(wc, w.nodes, w.cur);
    }
    float t[4];
    uint32_t ref[4];
    wide_test(w.nodes, w.cur, v3(w.o.x * w.inv.x, w.o.y * w.inv.y, w.o.z * w.inv.z), w.inv, w.tmin, w.best.t, t, ref);
    const int m = (t[0] < 3e38f) + (t[1] < 3e38f) + (t[2] < 3e38f) + (t[3] < 3e38f);
    float s0 = t[0], s1 = t[1], s2 = t[2], s3 = t[3];
    uint32_t c0 = ref[0], c1 = ref[1], c2 = ref[2], c3 = ref[3];
    cswap_(s0, c0, s1, c1);
    cswap_(s2, c2, s3, c3);
    cswap_(s0, c0, s2, c2);
    cswap_(s1, c1, s3, c3);
    cswap_(s1, c1, s2, c2);
    if (m == 0) {
        if (w.sp == 0) return true;
        w.sp--;
        w.cur = stk_get<STK>(S, w.sp);
        return false;
    }
    if (w.sp + m - 1 > S.cap) {   // cannot happen with a correct depth: said, never silently dropped
        atomicOr(overflow, 1u);
        return true;
    }
    auto push = [&](uint32_t x) {
        if (w.sp < STK) S.stk[w.sp * S.stride] = x;
        else S.spill[(size_t)(w.sp - STK) * S.sstride] = x;
        w.sp++;
    };
    if (m > 3) push(c3);
    if (m > 2) push(c2);
    if (m > 1) push(c1);
    w.cur = c0;
    return false;
}

// a wave's claimed queue entries not yet handed out, [rb, re), and whether the counter may still
// give entries
struct Refill {
    uint32_t rb = 0, re = 0;
    bool more = true;
};

// At the round loop's top (every lane of the wave there): the entry the calling lane takes
// (LR_NONE: none), idle lanes taking the wave's entries in lane order, a new run of 64 claimed
// when the wave's is too short.
__device__ __forceinline__ uint32_t refill_take(Refill &f, unsigned int *ctr, uint32_t n, bool idle_lane, uint64_t idle,
                                                uint32_t nidle)
{
    const uint32_t have = f.re - f.rb;
    uint32_t nb = 0, ne = 0;
    if (have < nidle && f.more) {
        uint32_t b = 0;
        if (lane_id() == 0) b = atomicAdd(ctr, 64u);
        b = __shfl(b, 0);
        if (b < n) {
            nb = b;
            ne = b + 64u < n ? b + 64u : n;
        }
        f.more = b + 64u < n;
    }
    uint32_t e = LR_NONE;
    if (idle_lane) {
        const uint32_t k = (uint32_t)__popcll(idle & lanemask_lt());
        if (k < have) e = f.rb + k;
        else if (k - have < ne - nb) e = nb + (k - have);
    }
    const uint32_t taken = nidle < have + (ne - nb) ? nidle : have + (ne - nb);
    if (taken <= have) f.rb += taken;
    else {
        f.rb = nb + (taken - have);
        f.re = ne;
    }
    return e;
}

// the extension walk with lanes refilled
template <int BLK, int STK, bool COUNT>
__global__ __launch_bounds__(BLK, walk_waves(BLK, STK)) void k_extend_rf(EngArgs A)
{
    __shared__ uint32_t stk[STK][BLK];
    EngCtl *ctl = A.ctl;
    const int par = (int)(ctl->iter & 1ull);
    const uint32_t n = ctl->count[par];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned long long c = ctl->claimed;
        ctl->head += c;
        ctl->next += c;
        ctl->claimed = 0;
        ctl->rays_ext += n;
        ctl->count[1 - par] = 0;
        ctl->count[2] = 0;
        reset_tiles(ctl, K_EXTEND);
    }
    const size_t gtid = (size_t)blockIdx.x * BLK + threadIdx.x;
    for (size_t k = gtid; k < (size_t)n / 64 + 2; k += (size_t)gridDim.x * BLK) A.ostat[k] = 0ull;
    const WalkStack S{&stk[0][threadIdx.x], BLK, A.spill + gtid, (size_t)A.spill_stride, STK + A.spill_per};
    WalkCnt wc;
    Refill f;
    uint32_t i = LR_NONE;   // the lane's queue entry (LR_NONE: idle)
    WalkState w;
    for (;;) {
        const uint64_t idle = __ballot(i == LR_NONE);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle == 64u && f.rb == f.re && !f.more) break;
        if (nidle >= (uint32_t)A.refill) {
            const uint32_t e = refill_take(f, &ctl->tiles[K_EXTEND], n, i == LR_NONE, idle, nidle);
            if (e != LR_NONE) {
                const JobDev &J = A.jobs[A.qid[par][e] >> LR_SLOT_BITS];
                const float4 r0 = A.qray[par][2 * (size_t)e], r1 = A.qray[par][2 * (size_t)e + 1];
                if (J.two_level) {   // not walked here: flagged, so that the host reports it
                    atomicOr(A.overflow, 2u);
                    A.hits[e] = make_float4(r1.w, __uint_as_float(LR_NONE), 0.0f, 0.0f);
                } else {
                    i = e;
                    walk_start(w, J, r0, r1);
                }
            }
        }
        if (i != LR_NONE && walk_round<false, COUNT, STK>(w, S, A.overflow, wc)) {
            A.hits[i] = make_float4(w.best.t, __uint_as_float(w.best.id), w.best.b1, w.best.b2);
            i = LR_NONE;
        }
    }
    if (COUNT) add_counts(A.counts, A.counts + 8, wc);
}

// the shadow walk with lanes refilled (single-level scenes, one shadow ray a vertex): an entry's
// light joins its path when the ray reaches it; paths that end are retired at the round's end,
// every lane of the wave there
template <int BLK, int STK, bool COUNT>
__global__ __launch_bounds__(BLK, walk_waves(BLK, STK)) void k_shadow_rf(EngArgs A)
{
    __shared__ uint32_t stk[STK][BLK];
    EngCtl *ctl = A.ctl;
    const uint32_t n = ctl->count[2];
    const int par = (int)ctl->spar;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        reset_tiles(ctl, K_SHADOW);
        ctl->rays_shadow += n;
        ctl->iter += 1;
    }
    const size_t gtid = (size_t)blockIdx.x * BLK + threadIdx.x;
    const WalkStack S{&stk[0][threadIdx.x], BLK, A.spill + gtid, (size_t)A.spill_stride, STK + A.spill_per};
    WalkCnt wc;
    Refill f;
    uint32_t i = LR_NONE;   // the lane's shadow entry (LR_NONE: idle)
    WalkState w;
    for (;;) {
        const uint64_t idle = __ballot(i == LR_NONE);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle == 64u && f.rb == f.re && !f.more) break;
        if (nidle >= (uint32_t)A.refill_shadow) {
            const uint32_t e = refill_take(f, &ctl->tiles[K_SHADOW], n, i == LR_NONE, idle, nidle);
            if (e != LR_NONE) {
                i = e;
                const JobDev &J = A.jobs[A.sid[e] >> LR_SLOT_BITS];
                walk_start(w, J, A.sray[2 * (size_t)e], A.sray[2 * (size_t)e + 1]);
                if (J.two_level) {   // not walked here: flagged, so that the host reports it (occluded)
                    atomicOr(A.overflow, 2u);
                    w.cur = LR_NONE;   // ends below at once, unlit
                    w.best.id = 0u;
                }
            }
        }
        bool ends = false, glow = false;
        uint32_t jslot = 0, q = 0;
        float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i != LR_NONE && (w.cur == LR_NONE || walk_round<true, COUNT, STK>(w, S, A.overflow, wc))) {
            const uint32_t id = A.sid[i];
            const uint32_t at = id & LR_SLOT_MASK;
            jslot = id >> LR_SLOT_BITS;
            glow = (at & 1u) != 0u;
            const float4 c = A.scon[i];
            ends = c.w != 0.0f;
            const bool lit = w.best.id == LR_NONE;
            if (lit || ends) {
                r = A.srad[i];
                if (lit) {
                    r.x += c.x;
                    r.y += c.y;
                    r.z += c.z;
                }
                if (!ends) A.rad[par][at] = r;   // the path's next state (shade wrote it without this)
                else q = A.sq[i];
            }
            i = LR_NONE;
        }
        if (__ballot(ends)) retire(A, ends, jslot, q, r, glow);
    }
    if (COUNT) add_counts(A.counts + 4, A.counts + 9, wc);
}

// The hit of a ray from o along d at t, shape hid (an instanced scene: inst is the instance):
// its outward normal and material (set_hit, intersect.c:4-10; an instance's normal by the
// inverse transpose, its material or its triangle's), ii the instance (LR_NONE: none)
__device__ __forceinline__ void hit_frame(const JobDev &J, V3 o, V3 d, float t, uint32_t hid, uint32_t inst, V3 hp,
                                          V3 &outward, int &mi, uint32_t &ii)
{
    ii = LR_NONE;
    if (J.two_level) {
        ii = inst;
        const float4 *rows = J.inst + 4 * (size_t)ii;
        const float4 info = rows[3];
        const uint32_t kind = __float_as_uint(info.x), imat = __float_as_uint(info.y);
        V3 oo, od, no;
        to_object(rows, o, d, oo, od);
        int tmat = 0;
        if (kind == LR_UNIT_SPHERE) no = vadd(oo, vscale(od, t));
        else if (kind == LR_UNIT_BOX) no = unit_box_normal(vadd(oo, vscale(od, t)));
        else {
            const float4 tu = J.tris[3 * (size_t)hid + 1], tv = J.tris[3 * (size_t)hid + 2];
            no = vcross(v3of(tu), v3of(tv));
            tmat = (int)(__float_as_uint(tv.w) >> 26);
        }
        outward = normal_to_world(rows, no);
        mi = imat != LR_NONE ? (int)imat : tmat;
    } else if (hid < J.nprims) {
        const Prim pr = load_prim(J.prims, hid);
        outward = pr.q.w < 0.5f ? vscale(vsub(hp, v3of(pr.c)), 1.0f / pr.c.w) : v3of(pr.n);
        mi = (int)pr.u.w;
    } else {
        const size_t ti = hid - J.nprims;
        const float4 tu = J.tris[3 * ti + 1], tv = J.tris[3 * ti + 2];
        outward = vnorm(vcross(v3of(tu), v3of(tv)));
        mi = (int)(__float_as_uint(tv.w) >> 26);
    }
}

// ---- glowing shapes as lights (Scene.emitters: tracer.c:120-216) ----
// 1 - cos of the half angle of the cone a ball of radius r fills from dist2 (squared) away, in
// the form without cancellation for a small far ball; 0 from inside (ball_cone)
__device__ __forceinline__ float ball_cone(float r, float dist2)
{
    const float sin2 = r * r / dist2;
    if (!(sin2 < 1.0f)) return 0.0f;
    return sin2 / (1.0f + sqrtf(1.0f - sin2));
}

// the density over solid angle (the pick included) NEE would have given emitter k's point at
// distance t along unit d from p, its outward normal n there (emit_pdf)
__device__ __forceinline__ float emit_pdf(const JobDev &J, int k, V3 p, V3 d, float t, V3 n)
{
    const float4 *e = J.emit + 5 * (size_t)k;
    const float4 e0 = e[0];
    if (e0.w != 0.0f) {
        const V3 to = vsub(v3of(e0), p);
        const float om = ball_cone(e[1].w, vdot(to, to));
        return om > 0.0f ? 1.0f / (2.0f * PI_F * om * (float)J.nemit) : 0.0f;
    }
    const float cl = -vdot(n, d);
    return cl > 0.0f ? t * t / (cl * e[2].w * (float)J.nemit) : 0.0f;
}

// a point of a glowing shape seen from p: direction, distance, density (the pick included) and
// radiance; false if none (p inside a ball, a box face turned away) (emit_sample)
__device__ __forceinline__ bool emit_sample(const JobDev &J, V3 p, float u0, float ua, float ub, V3 &wi, float &dist,
                                            float &pdf, V3 &le)
{
    int k = (int)(u0 * (float)J.nemit);
    if (k >= J.nemit) k = J.nemit - 1;
    const float4 *e = J.emit + 5 * (size_t)k;
    const float4 e0 = e[0], e1 = e[1], e2 = e[2], e3 = e[3];
    le = v3of(e[4]);
    const V3 c = v3of(e0);
    if (e0.w != 0.0f) {
        const float r = e1.w;
        const V3 to = vsub(c, p);
        const float d2 = vdot(to, to), om = ball_cone(r, d2);
        if (!(om > 0.0f)) return false;
        const float dc = sqrtf(d2), cos_t = 1.0f - ua * om, sin_t = sqrtf(rfmax(0.0f, 1.0f - cos_t * cos_t));
        const float ph = 2.0f * PI_F * ub;
        const V3 w = vscale(to, 1.0f / dc);
        V3 t, b;
        onb(w, t, b);
        wi = vadd(vscale(w, cos_t), vadd(vscale(t, sin_t * cosf(ph)), vscale(b, sin_t * sinf(ph))));
        const float disc = r * r - d2 * sin_t * sin_t;   // the ball's near side along wi
        dist = dc * cos_t - sqrtf(rfmax(0.0f, disc));
        pdf = 1.0f / (2.0f * PI_F * om * (float)J.nemit);
        return dist > 0.0f;
    }
    // a box: a face (two per axis, 2 a_j by 2 a_k) by its area, then a point on it (the six
    // faces walked in order 0+, 0-, 1+, 1-, 2+, 2-; selects, not arrays, keep it in registers)
    const V3 a0 = v3of(e1), a1 = v3of(e2), a2 = v3of(e3);
    const float fa0 = 4.0f * vlen(vcross(a1, a2)), fa1 = 4.0f * vlen(vcross(a2, a0)), fa2 = 4.0f * vlen(vcross(a0, a1));
    float x = ua * e2.w;
    int f = 0;
    for (; f < 5; f++) {
        const float faf = f < 2 ? fa0 : (f < 4 ? fa1 : fa2);
        if (!(x >= faf)) break;
        x -= faf;
    }
    const int i = f >> 1;
    const float fai = i == 0 ? fa0 : (i == 1 ? fa1 : fa2);
    const V3 ai = i == 0 ? a0 : (i == 1 ? a1 : a2), aj = i == 0 ? a1 : (i == 1 ? a2 : a0), ak = i == 0 ? a2 : (i == 1 ? a0 : a1);
    const float sg = f & 1 ? -1.0f : 1.0f, uu = rfmin(x / fai, 1.0f) * 2.0f - 1.0f, vv = ub * 2.0f - 1.0f;
    const V3 q = vadd(c, vadd(vscale(ai, sg), vadd(vscale(aj, uu), vscale(ak, vv))));
    V3 n = vscale(vnorm(vcross(aj, ak)), sg);
    if (vdot(n, ai) * sg < 0.0f) n = vneg(n);   // outward, whatever the axes' handedness
    const V3 to = vsub(q, p);
    const float d2 = vdot(to, to);
    dist = sqrtf(d2);
    wi = vscale(to, 1.0f / dist);
    const float cl = -vdot(n, wi);
    if (!(cl > 0.0f)) return false;
    pdf = d2 / (cl * e2.w * (float)J.nemit);
    return true;
}

// shade (wf_shade, trace_wave.wgsl:191-310): the bounce of trace_path_sampled (tracer.c:382-507)
// MULTI: the variant for scenes with more than one light (a lamp and a sky, glowing shapes): up
// to three shadow rays a vertex. Without it, one (the lamp's, or the sky's in a scene with no
// lamp), in fewer registers.
template <int BLK, bool MULTI>
__global__ __launch_bounds__(BLK) void k_shade(EngArgs A)
{
    EngCtl *ctl = A.ctl;
    const int cur = (int)(ctl->iter & 1ull);
    const uint32_t n = ctl->count[cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        reset_tiles(ctl, K_SHADE);
        ctl->spar = (unsigned int)(1 - cur);
    }
    for_tiles<BLK>(&ctl->tiles[K_SHADE], n, 1, [&](uint32_t i, bool live) {
        bool next = false, shadow = false;
        float4 nr0, nr1, sr0, sr1, scon;
        uint32_t jslot = 0, q = 0;
        int depth = 0, nsh = 0;
        bool glow = false;
        uint64_t st = 0;
        float4 rad_out = make_float4(0.0f, 0.0f, 0.0f, 0.0f), thr_out;
        if (live) {
            jslot = A.qid[cur][i] >> LR_SLOT_BITS;
            const JobDev &J = A.jobs[jslot];
            const uint2 mt = A.meta[cur][i];
            q = mt.x;
            depth = (int)(mt.y & 0x7fffffffu);
            glow = mt.y >> 31;   // the camera ray found a bright glow (glow_bright, trace.wgsl:1185)
            const float4 r0 = A.qray[cur][2 * (size_t)i], r1 = A.qray[cur][2 * (size_t)i + 1];
            const float4 hv = A.hits[i];
            const float4 thr4 = A.thr[cur][i], rad4 = A.rad[cur][i];
            V3 thr = v3of(thr4), rad = v3of(rad4);
            float bounce_pdf = thr4.w;
            // the emission state (low byte), the caustic chain (bit 8: the camera saw a rough surface
            // first, light tracing on) and whether the last vertex was a mirror or glass (bit 9)
            // (tracer.c:371-375)
            int emit = __float_as_int(rad4.w);
            const uint32_t hid = __float_as_uint(hv.y);
            const V3 d = v3of(r1), o = v3of(r0);
            const uint32_t pixel = q % (uint32_t)J.npix;
            const uint32_t k = J.s0 + q / (uint32_t)J.npix;
            st = A.rng[cur][i];
            const uint64_t inc = ((uint64_t)pixel << 1u) | 1u;
            const uint64_t seed = ((uint64_t)J.seed_hi << 32) | J.seed_lo;
            const Smp sm{(int)(pixel % (uint32_t)J.W), (int)(pixel / (uint32_t)J.W), k, (uint32_t)seed * 0x9e3779b9u,
                         A.mask, &st, inc};
            const EnvRef env{J.env_rgb, J.env_marg, J.env_cond, J.env_prob, J.env_w, J.env_h, J.env_scale, J.env_rot,
                             J.env_rot_c, J.env_rot_s};
            if (hid == LR_NONE) {
                // escaped: the sky (tracer.c:415-421), black without one
                if (J.env_w > 0 && (emit & 0xff) != EMIT_NONE) {
                    V3 le = env_eval(env, d);
                    if ((emit & 0xff) == EMIT_MIS) le = vscale(le, power_heuristic(bounce_pdf, env_pdf(env, d)));
                    rad = vadd(rad, vmul(thr, le));
                }
            } else {
                // the hit (set_hit, intersect.c:4-10): point, the normal facing the ray
                const float t = hv.x;
                const V3 hp = vadd(o, vscale(d, t));
                V3 outward;
                int mi;
                uint32_t ii;
                hit_frame(J, o, d, t, hid, __float_as_uint(hv.z), hp, outward, mi, ii);
                const bool front = vdot(d, outward) < 0.0f;
                const V3 nrm = front ? outward : vneg(outward);
                int type;
                Bsdf b = bsdf_make(J.mats, mi, front, A.tab, type);
                if (depth == 0 && type != MAT_LIGHT) {   // (a glow, not a lamp: trace.wgsl:1183-1186)
                    const float4 e1 = J.mats[3 * mi + 1];
                    if (rfmax(e1.x, rfmax(e1.y, e1.z)) >= 1.0f) glow = true;
                }
                if (type == MAT_LIGHT) {
                    if (front && (emit & 0x300) != 0x300) {   // (camera -> rough ... mirror/glass -> lamp: light tracing's)
                        const V3 em = v3of(J.mats[3 * mi + 1]);
                        if ((emit & 0xff) == EMIT_FULL) rad = vadd(rad, vmul(thr, em));
                        else if ((emit & 0xff) == EMIT_MIS && (int)hid == J.light) {
                            const Prim l = load_prim(J.prims, (uint32_t)J.light);
                            const float cos_light = -vdot(v3of(l.n), d);
                            if (cos_light > 0.0f) {
                                const float w = power_heuristic(bounce_pdf, (t * t) / (cos_light * l.v.w));
                                rad = vadd(rad, vmul(thr, vscale(em, w)));
                            }
                        }
                    }
                } else {
                    // a glowing surface (material_glows: not a lamp, some emission; tracer.c:450-458):
                    // found by the path; a glowing shape reached after a rough bounce was sampled by
                    // NEE too, so weighted against that (MIS) or, without MIS, counted there already
                    const V3 em = v3of(J.mats[3 * mi + 1]);
                    if (em.x > 0.0f || em.y > 0.0f || em.z > 0.0f) {
                        const int k = MULTI && J.nemit && (emit & 0xff) != EMIT_FULL && ii != LR_NONE ? J.inst_emit[ii] : -1;
                        const float pe = k >= 0 && front ? emit_pdf(J, k, o, vnorm(d), t * vlen(d), nrm) : 0.0f;
                        if (!(pe > 0.0f)) rad = vadd(rad, vmul(thr, em));
                        else if ((emit & 0xff) == EMIT_MIS) rad = vadd(rad, vmul(thr, vscale(em, power_heuristic(bounce_pdf, pe))));
                    }
                }
                if (type != MAT_LIGHT && depth < J.max_depth) {
                    const V3 v = vneg(vnorm(d));
                    // 1. next event estimation of the lamp, the sky and the glowing shapes, in that
                    //    order (tracer.c:467-478): each a shadow ray whose visibility the shadow
                    //    kernel decides, adding the contributions in this order. The first ray
                    //    waits for its place in the shadow queue; the others go beside entry i.
                    const bool rough_part = !b.smooth || (1.0f - b.metallic) * (1.0f - b.transmission) > 0.0f;
                    auto add_shadow = [&](V3 wi, float tmax, V3 c) {
                        const float4 a = make_float4(hp.x, hp.y, hp.z, J.eps), w = make_float4(wi.x, wi.y, wi.z, tmax);
                        const float4 cc = make_float4(c.x, c.y, c.z, 0.0f);
                        if (!MULTI || nsh == 0) {
                            sr0 = a;
                            sr1 = w;
                            scon = cc;
                        } else {
                            float4 *xr = nsh == 1 ? A.xsray[0] : A.xsray[1];   // (selects: no indexed copy of the arguments)
                            xr[2 * (size_t)i] = a;
                            xr[2 * (size_t)i + 1] = w;
                            (nsh == 1 ? A.xscon[0] : A.xscon[1])[i] = cc;
                        }
                        nsh++;
                    };
                    // the sky (sample_direct_env, tracer.c:87-101): the lamp's pair when there is no
                    // lamp, else two plain numbers (env_numbers, :232-235)
                    auto sky_nee = [&]() {
                        float ua, ub;
                        if (!MULTI || J.light < 0) s2(sm, 1 + 3 * depth, ua, ub);
                        else {
                            ua = pcg_float(st, inc);
                            ub = pcg_float(st, inc);
                        }
                        V3 wi;
                        float pdf;
                        const V3 le = env_sample(env, ua, ub, wi, pdf);
                        if (pdf > 0.0f && vmax3(le) > 0.0f) {
                            float bpdf;
                            const V3 f = bsdf_eval(b, A.tab, nrm, v, wi, bpdf);
                            if (vmax3(f) > 0.0f) {
                                float g = fabsf(vdot(nrm, wi)) / pdf;
                                if (J.mis) g *= power_heuristic(pdf, bpdf);
                                add_shadow(wi, 1e30f, vmul(thr, vscale(vmul(f, le), g)));
                            }
                        }
                    };
                    // (one light: the sky's block first, which the compiler allocates fewer registers for)
                    if constexpr (!MULTI)
                        if (rough_part && J.env_w > 0 && J.light < 0) sky_nee();
                    if (rough_part && J.light >= 0) {
                        // the lamp (sample_direct_light, tracer.c:56-81)
                        float la, lb;
                        s2(sm, 1 + 3 * depth, la, lb);
                        const float ua = lb, ub = la;   // the pair's second number first (tracer.c:50-53)
                        const Prim l = load_prim(J.prims, (uint32_t)J.light);
                        const V3 le = v3of(J.mats[3 * (int)l.u.w + 1]);
                        const V3 xl = vadd(v3of(l.q), vadd(vscale(v3of(l.u), ua), vscale(v3of(l.v), ub)));
                        const V3 to_light = vsub(xl, hp);
                        const float dist2 = vdot(to_light, to_light);
                        const float dist = sqrtf(dist2);
                        const V3 wi = vscale(to_light, 1.0f / dist);
                        const float cos_surf = vdot(nrm, wi);
                        const float cos_light = -vdot(v3of(l.n), wi);
                        if (cos_light > 0.0f) {
                            float bpdf;
                            const V3 f = bsdf_eval(b, A.tab, nrm, v, wi, bpdf);
                            if (vmax3(f) > 0.0f) {
                                float g = fabsf(cos_surf) * cos_light * l.v.w / dist2;
                                if (J.mis) g *= power_heuristic(dist2 / (cos_light * l.v.w), bpdf);
                                add_shadow(wi, dist - J.margin, vmul(thr, vscale(vmul(f, le), g)));
                            }
                        }
                    }
                    if constexpr (MULTI)
                        if (rough_part && J.env_w > 0) sky_nee();
                    if (MULTI && rough_part && J.nemit > 0) {
                        // the glowing shapes (sample_direct_emit, tracer.c:202-216): three plain numbers
                        const float u0 = pcg_float(st, inc), ua = pcg_float(st, inc), ub = pcg_float(st, inc);
                        V3 wi, le;
                        float dist, pdf;
                        if (emit_sample(J, hp, u0, ua, ub, wi, dist, pdf, le)) {
                            float bpdf;
                            const V3 f = bsdf_eval(b, A.tab, nrm, v, wi, bpdf);
                            if (vmax3(f) > 0.0f) {
                                float g = fabsf(vdot(nrm, wi)) / pdf;
                                if (J.mis) g *= power_heuristic(pdf, bpdf);
                                add_shadow(wi, dist - J.margin, vmul(thr, vscale(vmul(f, le), g)));
                            }
                        }
                    }
                    shadow = nsh > 0;
                    // 2. the next direction from the BSDF, Russian roulette from depth 3
                    bsdf_view(b, A.tab, vdot(v, nrm));
                    float ur = 0.0f, ul = 0.0f, da = 0.0f, db = 0.0f;
                    const int nl = (b.p[0] > 0.0f) + (b.p[1] > 0.0f) + (b.p[2] > 0.0f) + (b.p[3] > 0.0f);
                    if (depth >= 3 || nl > 1) s2(sm, 3 + 3 * depth, ur, ul);
                    const int lobe = bsdf_pick(b, ul);
                    if (lobe >= 0 && !(b.smooth && (lobe == SPEC || lobe == TRANS))) s2(sm, 2 + 3 * depth, da, db);
                    BSample bs;
                    if (bsdf_sample_lobe(b, A.tab, lobe, nrm, v, da, db, bs)) {
                        thr = vmul(thr, bs.w);
                        if (bs.delta) emit = EMIT_FULL | (emit & 0x100) | 0x200;
                        else {
                            bounce_pdf = bs.pdf;
                            emit = (J.mis ? EMIT_MIS : EMIT_NONE) | (depth == 0 ? (J.skip_caustics ? 0x100 : 0) : (emit & 0x100));
                        }
                        bool go = true;
                        if (depth >= 3) {
                            const float pr = rfmin(vmax3(thr), 0.95f);
                            if (ur >= pr) go = false;
                            else thr = vscale(thr, 1.0f / pr);
                        }
                        if (go) {
                            const V3 nd = J.mis ? vnorm(bs.l) : bs.l;
                            next = true;
                            nr0 = make_float4(hp.x, hp.y, hp.z, J.eps);
                            nr1 = make_float4(nd.x, nd.y, nd.z, 1e30f);
                        }
                    }
                }
            }
            rad_out = make_float4(rad.x, rad.y, rad.z, __int_as_float(emit));
            thr_out = make_float4(thr.x, thr.y, thr.z, bounce_pdf);
            if (shadow && !next) scon.w = 1.0f;   // the shadow kernel ends the path
        }
        uint32_t at, ss;
        ordered_append2<BLK>(A.ostat, i / BLK, (n + BLK - 1) / BLK, next, shadow, at, ss, &ctl->count[1 - cur],
                             &ctl->count[2]);
        // the surviving path: its next ray and its state, side by side in the other queue
        if (next) {
            A.qray[1 - cur][2 * (size_t)at] = nr0;
            A.qray[1 - cur][2 * (size_t)at + 1] = nr1;
            A.qid[1 - cur][at] = jslot << LR_SLOT_BITS;
            A.thr[1 - cur][at] = thr_out;
            A.rad[1 - cur][at] = rad_out;
            A.rng[1 - cur][at] = st;
            A.meta[1 - cur][at] = make_uint2(q, ((uint32_t)depth + 1u) | (glow ? 0x80000000u : 0u));
        }
        // the shadow ray carries the path's radiance and q (its contribution is added to one or
        // the other: the path's next state, or its pixel if the path ends with it)
        if (shadow) {
            A.sray[2 * (size_t)ss] = sr0;
            A.sray[2 * (size_t)ss + 1] = sr1;
            A.sid[ss] = (jslot << LR_SLOT_BITS) | (next ? at : (glow ? 1u : 0u));   // (a path that ends: its glow flag)
            A.scon[ss] = scon;
            A.srad[ss] = rad_out;
            A.sq[ss] = q;
            if (MULTI) A.sx[ss] = ((uint32_t)(nsh - 1) << 30) | i;
        }
        retire(A, live && !next && !shadow, jslot, q, rad_out, glow);
    });
}

// shadow (wf_shadow, trace_wave.wgsl:335-355): any hit before the lamp; if none, the
// contribution joins the path's radiance. Paths whose last vertex made the shadow ray end here.
// TL and the register bound as for extend.
template <int BLK, int STK, bool COUNT, bool TL, bool MULTI>
__global__ __launch_bounds__(BLK, TL ? 1 : walk_waves(BLK, STK)) void k_shadow(EngArgs A)
{
    __shared__ uint32_t stk[STK][BLK];
    EngCtl *ctl = A.ctl;
    const uint32_t n = ctl->count[2];
    const int par = (int)ctl->spar;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        reset_tiles(ctl, K_SHADOW);
        ctl->rays_shadow += n;
        ctl->iter += 1;
    }
    const size_t gtid = (size_t)blockIdx.x * BLK + threadIdx.x;
    const WalkStack S{&stk[0][threadIdx.x], BLK, A.spill + gtid, (size_t)A.spill_stride, STK + A.spill_per};
    WalkCnt wc;
    for_tiles<BLK>(&ctl->tiles[K_SHADOW], n, A.mult, [&](uint32_t i, bool live) {
        bool ends = false;
        uint32_t jslot = 0, q = 0;
        float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int nx = 0;
        bool glow = false;
        if (live) {
            const uint32_t id = A.sid[i];
            const uint32_t at = id & LR_SLOT_MASK;
            jslot = id >> LR_SLOT_BITS;
            glow = (at & 1u) != 0u;   // (read only for a path that ends here: then at holds its glow flag)
            const JobDev &J = A.jobs[jslot];
            if constexpr (!MULTI) {
                const float4 r0 = A.sray[2 * (size_t)i], r1 = A.sray[2 * (size_t)i + 1];
                WalkHit h;
                if (TL && J.two_level) h = walk4_tl<true, COUNT, STK>(J, v3of(r0), v3of(r1), r0.w, r1.w, S, A.overflow, wc);
                else if (!J.two_level)
                    h = walk4<true, COUNT, STK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, v3of(r0), v3of(r1), r0.w, r1.w, S,
                                                A.overflow, wc);
                else {
                    atomicOr(A.overflow, 2u);
                    h = WalkHit{0.0f, 0u, 0.0f, 0.0f};   // occluded: adds nothing
                }
                const float4 c = A.scon[i];
                ends = c.w != 0.0f;
                const bool lit = h.id == LR_NONE;
                if (lit || ends) {
                    r = A.srad[i];
                    if (lit) {
                        r.x += c.x;
                        r.y += c.y;
                        r.z += c.z;
                    }
                    if (!ends) A.rad[par][at] = r;   // the path's next state (shade wrote it without this)
                    else q = A.sq[i];
                }
            } else {
                const uint32_t x = A.sx[i];
                nx = (int)(x >> 30);                       // the vertex's rays past the first
                const uint32_t src = x & LR_SLOT_MASK;     // where shade left them
                const float4 c0 = A.scon[i];
                ends = c0.w != 0.0f;
                uint32_t lit = 0;                          // bit j: ray j reached its light
                for (int j = 0; j <= nx; j++) {
                    const float4 *rr = j == 0 ? A.sray + 2 * (size_t)i : (j == 1 ? A.xsray[0] : A.xsray[1]) + 2 * (size_t)src;
                    const float4 r0 = rr[0], r1 = rr[1];
                    WalkHit h;
                    if (TL && J.two_level) h = walk4_tl<true, COUNT, STK>(J, v3of(r0), v3of(r1), r0.w, r1.w, S, A.overflow, wc);
                    else if (!J.two_level)
                        h = walk4<true, COUNT, STK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, v3of(r0), v3of(r1), r0.w, r1.w, S,
                                                    A.overflow, wc);
                    else {
                        atomicOr(A.overflow, 2u);
                        h = WalkHit{0.0f, 0u, 0.0f, 0.0f};   // occluded: adds nothing
                    }
                    if (h.id == LR_NONE) lit |= 1u << j;
                }
                if (lit || ends) {
                    // the lights' contributions join the path's radiance in the order shade made them
                    r = A.srad[i];
                    for (int j = 0; j <= nx; j++)
                        if (lit >> j & 1u) {
                            const float4 c = j == 0 ? c0 : (j == 1 ? A.xscon[0] : A.xscon[1])[src];
                            r.x += c.x;
                            r.y += c.y;
                            r.z += c.z;
                        }
                    if (!ends) A.rad[par][at] = r;   // the path's next state (shade wrote it without these)
                    else q = A.sq[i];
                }
            }
        }
        // the rays past an entry's first are shadow rays too: counted a wave at a time
        if constexpr (MULTI) {
            if (__ballot(nx > 0)) {
                int cnt = nx;
                for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
                if ((threadIdx.x & 63) == 0) atomicAdd(&ctl->rays_shadow, (unsigned long long)cnt);
            }
        }
        retire(A, ends, jslot, q, r, glow);
    });
    if (COUNT) add_counts(A.counts + 4, A.counts + 9, wc);
}

// The denoiser's guide of a job (trace.wgsl:1034-1107, guide): each pixel's ray through its
// centre, through mirrors and glass (refracted unless the glass reflects all) to the first
// surface light scatters on: its normal facing the ray, its distance along the way, and the
// albedo of the whole way to it (kind 0); a lamp, a surface glowing at least as bright as white
// paper under a light of 1 (glow_bright, :1115), or nothing within 6 vertices: kind 1, the
// albedo 1. A grid-stride loop over the job's pixels; the walks' stacks spill as extend's do.
template <int BLK, int STK, bool TL>
__global__ __launch_bounds__(BLK) void k_guide(JobDev J, float4 *guide, uint32_t *spill, size_t spill_stride, int spill_per,
                                              const float *tab, uint32_t *overflow)
{
    __shared__ uint32_t stk[STK][BLK];
    const size_t gtid = (size_t)blockIdx.x * BLK + threadIdx.x;
    const WalkStack S{&stk[0][threadIdx.x], BLK, spill + gtid, spill_stride, STK + spill_per};
    for (size_t p = gtid; p < (size_t)J.npix; p += (size_t)gridDim.x * BLK) {
        const int x = (int)(p % (uint32_t)J.W), y = (int)(p / (uint32_t)J.W);
        const float u = ((float)x + 0.5f) / (float)J.W, v = ((float)y + 0.5f) / (float)J.H;
        const V3 org = v3(J.cam[0], J.cam[1], J.cam[2]);
        const V3 ul = v3(J.cam[3], J.cam[4], J.cam[5]), hz = v3(J.cam[6], J.cam[7], J.cam[8]),
                 vt = v3(J.cam[9], J.cam[10], J.cam[11]);
        V3 o = org, d = vnorm(vsub(vsub(vadd(ul, vscale(hz, u)), vscale(vt, v)), org));
        V3 gn = v3(0.0f, 0.0f, 0.0f), ga = v3(1.0f, 1.0f, 1.0f), a = v3(1.0f, 1.0f, 1.0f);
        float gz = 1e30f, kind = 1.0f, dist = 0.0f;
        for (int i = 0; i < 6; i++) {
            WalkCnt wc;
            WalkHit h;
            if (TL && J.two_level) h = walk4_tl<false, false, STK>(J, o, d, J.eps, 1e30f, S, overflow, wc);
            else h = walk4<false, false, STK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, o, d, J.eps, 1e30f, S, overflow, wc);
            if (h.id == LR_NONE) break;
            const V3 hp = vadd(o, vscale(d, h.t));
            V3 outward;
            int mi;
            uint32_t ii;
            hit_frame(J, o, d, h.t, h.id, __float_as_uint(h.b1), hp, outward, mi, ii);
            const bool front = vdot(d, outward) < 0.0f;
            const V3 nrm = front ? outward : vneg(outward);
            const float4 m0 = J.mats[3 * mi], m1 = J.mats[3 * mi + 1];
            int type;
            const Bsdf b = bsdf_make(J.mats, mi, front, tab, type);
            if (type == MAT_LIGHT || rfmax(m1.x, rfmax(m1.y, m1.z)) >= 1.0f) {
                gn = nrm;
                gz = dist + h.t;
                break;
            }
            if (!b.smooth || (1.0f - b.metallic) * (1.0f - b.transmission) > 0.0f) {   // bsdf_rough_part
                gn = nrm;
                gz = dist + h.t;
                ga = vmul(a, v3of(m0));
                kind = 0.0f;
                break;
            }
            a = vmul(a, v3of(m0));
            dist += h.t;
            V3 nxt = vreflect(d, nrm);
            if (b.transmission > 0.5f && b.metallic < 0.5f) {
                const float cos_i = rfmin(vdot(vneg(d), nrm), 1.0f);
                if (b.eta * b.eta * (1.0f - cos_i * cos_i) <= 1.0f) nxt = vrefract(d, nrm, b.eta);
            }
            o = hp;
            d = vnorm(nxt);
        }
        guide[2 * p] = make_float4(gn.x, gn.y, gn.z, gz);
        guide[2 * p + 1] = make_float4(ga.x, ga.y, ga.z, kind);
    }
}

// ---- light tracing for caustics (light_trace_caustics, tracer.c:511-665; ltrace, trace.wgsl:1341-1511) ----
// A light path (pass k, path i of a pass's nlp; PCG32 stream i, sample index 2^31 + s0 + k): from
// a uniform point of the lamp in a cosine-distributed direction about its normal, power Le area
// pi; at every rough vertex after the first it is joined to the camera by a shadow ray and added
// to the pixel it projects to, in the job's image, over nlp (each pass's nlp paths are one
// estimate, added to that pass's sample of the pixels); it goes on by sampling the BSDF seen from
// the light's side (through a surface times eta^2, the adjoint of the radiance's factor), Russian
// roulette from depth 3. The first vertex must be a mirror or glass (lamp -> rough is the path
// tracer's, by NEE), which leaves exactly the paths the path tracer skips (lamp -> mirror/glass
// ... -> rough -> camera). No smooth normals here: the shading and geometric normals are one,
// the adjoint correction 1.
// Most light paths end at their first vertex (a rough surface). So the paths are traced in two
// launches: every path's emission and first vertex, the few that go on appended (their state and
// generator) to a queue; then those from their second vertex on, one a lane, so that a wave's lanes
// are not held by the one path in 64 that goes on. The numbers are drawn in the same order either
// way and the splats are integer sums: the image is the same.
struct LightState {
    V3 o, d, beta, thr;
    uint64_t st;
    uint32_t path;
};
struct LightCtx {
    Prim lamp;
    V3 le, org, hz, vt, fw;
    float pix_area, inv_n;
    uint64_t seed;
};
__device__ __forceinline__ LightCtx light_ctx(const JobDev &J, uint32_t nlp)
{
    LightCtx L;
    L.lamp = load_prim(J.prims, (uint32_t)J.light);
    L.le = v3of(J.mats[3 * (int)L.lamp.u.w + 1]);
    L.org = v3(J.cam[0], J.cam[1], J.cam[2]);
    const V3 ul = v3(J.cam[3], J.cam[4], J.cam[5]);
    L.hz = v3(J.cam[6], J.cam[7], J.cam[8]);
    L.vt = v3(J.cam[9], J.cam[10], J.cam[11]);
    L.pix_area = vlen(L.hz) * vlen(L.vt) / ((float)J.W * (float)J.H);   // camera_screen_area / pixels
    const V3 center = vadd(ul, vsub(vscale(L.hz, 0.5f), vscale(L.vt, 0.5f)));
    L.fw = vsub(center, L.org);
    L.inv_n = 1.0f / (float)nlp;
    L.seed = ((uint64_t)J.seed_hi << 32) | J.seed_lo;
    return L;
}

// ---- the photon cache (gpu_trace_photons, trace_gpu.c:1964-1996; pc_claim / deposit,
// trace.wgsl:1359-1404): a world-space hash grid of cells in three sizes (the finest, x4, x16),
// one set per main axis of the surface's normal and its sign. A cell's key (its integer corner,
// 19 bits an axis, the axis 3 bits, the size 2 bits) mixed by SplitMix64's finaliser (Steele, Lea
// and Flood 2014) gives its home slot (low bits) and its tag (high 32 bits, never 0); a cell takes
// the first slot from its home on that is free or its own, at most 8 on (linear probing), by a
// compare-and-swap of the tag. Sums are integers: any order of deposits gives the same table.
__device__ __forceinline__ uint64_t pc_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t pc_key(int x, int y, int z, int axis, int level)
{
    const uint64_t m = (1ull << 19) - 1ull;
    return pc_mix(((uint64_t)(uint32_t)x & m) | (((uint64_t)(uint32_t)y & m) << 19) | (((uint64_t)(uint32_t)z & m) << 38) |
                  ((uint64_t)axis << 57) | ((uint64_t)level << 60));
}
// the main axis of n and its sign: 2 axis + (n[axis] < 0)
__device__ __forceinline__ int pc_axis(V3 n)
{
    const float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    const int a = ax >= ay && ax >= az ? 0 : (ay >= az ? 1 : 2);
    const float v = a == 0 ? n.x : (a == 1 ? n.y : n.z);
    return 2 * a + (v < 0.0f ? 1 : 0);
}
__device__ __forceinline__ uint32_t pc_tag(uint64_t h) { return (uint32_t)(h >> 32) | 1u; }
// the cell's slot (claimed if need be), or LR_NONE when the 8 slots from its home are others'. A
// slot's tag only goes from 0 to a cell's during a trace, so a plain read that shows a tag is the
// answer; only a slot that reads free takes the compare-and-swap
__device__ __forceinline__ uint32_t pc_claim(unsigned long long *pc, uint64_t h)
{
    const uint32_t tag = pc_tag(h);
    uint32_t idx = (uint32_t)h & (LR_PC_SLOTS - 1u);
    for (int probe = 0; probe < 8; probe++) {
        uint32_t *w = reinterpret_cast<uint32_t *>(pc + 4 * (size_t)idx) + 7;
        uint32_t old = __atomic_load_n(w, __ATOMIC_RELAXED);
        if (old == 0u) old = atomicCAS(w, 0u, tag);
        if (old == 0u || old == tag) return idx;
        idx = (idx + 1u) & (LR_PC_SLOTS - 1u);
    }
    return LR_NONE;
}
// a point's cell of a size: along the normal's main axis offset by 1/pi of a cell, so that an
// axis-aligned surface at a round coordinate (a floor at y = 0) does not lie on the cells'
// boundary, where rounding would put some of its points in the layer of cells on either side
constexpr float PC_OFFSET = 0.31830988f;
__device__ __forceinline__ int pc_coord(float x, float size, bool normal_axis)
{
    return (int)floorf(x / size + (normal_axis ? PC_OFFSET : 0.0f));
}
// a work-group's sums before they go to the table: PC_LDS cells keyed by their slot (open
// addressing, 8 probes; a photon whose cell finds no room goes to the table at once). A focused
// caustic sends many photons to a few cells, whose adds would otherwise queue at their slots.
constexpr int PC_LDS = 256;
struct PcLds {
    unsigned long long sum[PC_LDS][3];
    uint32_t cnt[PC_LDS];
    uint32_t slot[PC_LDS];
};
__device__ __forceinline__ void pc_lds_init(PcLds &L)
{
    for (int i = threadIdx.x; i < PC_LDS; i += blockDim.x) {
        L.sum[i][0] = L.sum[i][1] = L.sum[i][2] = 0ull;
        L.cnt[i] = 0u;
        L.slot[i] = LR_NONE;
    }
}
__device__ __forceinline__ void pc_lds_flush(const JobDev &J, const PcLds &L)
{
    for (int i = threadIdx.x; i < PC_LDS; i += blockDim.x) {
        if (L.slot[i] == LR_NONE) continue;
        unsigned long long *w = J.pc + 4 * (size_t)L.slot[i];
        for (int ch = 0; ch < 3; ch++)
            if (L.sum[i][ch]) atomicAdd(&w[ch], L.sum[i][ch]);
        atomicAdd(reinterpret_cast<uint32_t *>(w + 3), L.cnt[i]);
    }
}
// a photon of power c at p on a surface facing n: into its cell of each size (through the
// work-group's sums L when given)
__device__ __forceinline__ void pc_deposit(const JobDev &J, V3 p, V3 n, V3 c, PcLds *L)
{
    const int axis = pc_axis(n), ax = axis / 2;
    float size = J.pc_cell;
    for (int level = 0; level < 3; level++, size *= 4.0f) {
        const uint32_t slot = pc_claim(J.pc, pc_key(pc_coord(p.x, size, ax == 0), pc_coord(p.y, size, ax == 1),
                                                    pc_coord(p.z, size, ax == 2), axis, level));
        if (slot == LR_NONE) continue;
        const float cc[3] = {c.x, c.y, c.z};
        unsigned long long v[3];
        for (int ch = 0; ch < 3; ch++) {
            const float x = cc[ch] * LR_PC_SCALE;
            v[ch] = (x > 0.0f && x <= 1.8e18f) ? (unsigned long long)__float2ull_rn(x) : 0ull;
        }
        int at = -1;
        if (L) {
            uint32_t i = slot & (PC_LDS - 1);
            for (int probe = 0; probe < 8; probe++, i = (i + 1u) & (PC_LDS - 1)) {
                const uint32_t old = atomicCAS(&L->slot[i], LR_NONE, slot);
                if (old == LR_NONE || old == slot) {
                    at = (int)i;
                    break;
                }
            }
        }
        if (at >= 0) {
            for (int ch = 0; ch < 3; ch++)
                if (v[ch]) atomicAdd(&L->sum[at][ch], v[ch]);
            atomicAdd(&L->cnt[at], LR_PC_UNIT);
        } else {
            unsigned long long *w = J.pc + 4 * (size_t)slot;
            for (int ch = 0; ch < 3; ch++)
                if (v[ch]) atomicAdd(&w[ch], v[ch]);
            atomicAdd(reinterpret_cast<uint32_t *>(w + 3), LR_PC_UNIT);
        }
    }
}

// one vertex of a light path at `depth`: the walk to it, its rough part joined to the camera
// (after the first vertex; PC: left in the photon cache instead), the next direction drawn;
// false: the path ends here
template <int STK, bool COUNT, bool PC = false>
__device__ __forceinline__ bool light_vertex(const JobDev &J, const LightCtx &L, int depth, LightState &P, const WalkStack &S,
                                             const float *tab, uint32_t *overflow, WalkCnt &wc, uint32_t &n_ext,
                                             uint32_t &n_shadow, PcLds *pl)
{
    const uint64_t inc = ((uint64_t)P.path << 1u) | 1u;
    const WalkHit h = walk4<false, COUNT, STK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, P.o, P.d, J.eps, 1e30f, S,
                                               overflow, wc);
    n_ext++;
    if (h.id == LR_NONE) return false;
    const V3 hp = vadd(P.o, vscale(P.d, h.t));
    V3 outward;
    int mi;
    uint32_t ii;
    hit_frame(J, P.o, P.d, h.t, h.id, __float_as_uint(h.b1), hp, outward, mi, ii);
    const bool front = vdot(P.d, outward) < 0.0f;
    const V3 nrm = front ? outward : vneg(outward);
    int type;
    Bsdf b = bsdf_make(J.mats, mi, front, tab, type);
    if (type == MAT_LIGHT) return false;
    const V3 wi = vneg(P.d);
    // 1. the rough part joined to the camera, or left in the photon cache (not at the first vertex)
    if (PC && depth > 0 && (!b.smooth || (1.0f - b.metallic) * (1.0f - b.transmission) > 0.0f))
        pc_deposit(J, hp, nrm, P.beta, pl);
    if (!PC && depth > 0 && (!b.smooth || (1.0f - b.metallic) * (1.0f - b.transmission) > 0.0f)) {
        const V3 to_cam = vsub(L.org, hp);
        const float dist2 = vdot(to_cam, to_cam), dist = sqrtf(dist2);
        const V3 wo = vscale(to_cam, 1.0f / dist);
        const float cos_o = vdot(nrm, wo);
        // the BSDF seen from the camera: its side's normal and index
        Bsdf bc = b;
        int tc;
        if (!(cos_o > 0.0f)) bc = bsdf_make(J.mats, mi, !front, tab, tc);
        const V3 nc = cos_o > 0.0f ? nrm : vneg(nrm);
        float pdf;
        V3 f = v3(0.0f, 0.0f, 0.0f);
        if (cos_o != 0.0f && vdot(nc, wo) > 0.0f) f = bsdf_eval(bc, tab, nc, wo, wi, pdf);
        // camera_project (camera.h:48-59)
        const V3 qv = vsub(hp, L.org);
        const float z = vdot(qv, L.fw);
        if (vmax3(f) > 0.0f && z > 0.0f) {
            const V3 e = vsub(vscale(qv, 1.0f / z), L.fw);
            const float sx = 0.5f + vdot(e, L.hz) / vdot(L.hz, L.hz), sy = 0.5f - vdot(e, L.vt) / vdot(L.vt, L.vt);
            const float cos_c = z / vlen(qv);
            if (sx >= 0.0f && sy >= 0.0f && sx < 1.0f && sy < 1.0f) {
                const int px = (int)(sx * (float)J.W), py = (int)(sy * (float)J.H);
                if (px < J.W && py < J.H) {
                    const WalkHit sh = walk4<true, COUNT, STK>(J.nodes, J.aux, J.ltris, J.prims, J.nprims, hp, wo, J.eps,
                                                               dist * (1.0f - 1e-4f), S, overflow, wc);
                    n_shadow++;
                    if (sh.id == LR_NONE) {
                        const float wgt = fabsf(cos_o) / (dist2 * L.pix_area * cos_c * cos_c * cos_c);
                        const V3 c = vscale(vmul(P.beta, f), wgt * L.inv_n);
                        long long *px3 = J.accum + 3 * ((size_t)py * J.W + px);
                        const float cc[3] = {c.x, c.y, c.z};
                        for (int ch = 0; ch < 3; ch++) {
                            if (!(cc[ch] > 0.0f)) continue;   // NaN or nothing
                            const long long v = to_fixed(cc[ch]);
                            if (v) atomicAdd((unsigned long long *)&px3[ch], (unsigned long long)v);
                        }
                    }
                }
            }
        }
    }
    // 2. on, by the BSDF seen from the light's side (bsdf_sample, brdf.h:343-348)
    const float ulb = pcg_float(P.st, inc), u1 = pcg_float(P.st, inc), u2 = pcg_float(P.st, inc);
    const float cv = vdot(wi, nrm);
    if (!(cv > 0.0f)) return false;
    bsdf_view(b, tab, cv);
    BSample bs;
    if (!bsdf_sample_lobe(b, tab, bsdf_pick(b, ulb), nrm, wi, u1, u2, bs)) return false;
    if (depth == 0 && !bs.delta) return false;   // lamp -> rough: the path tracer's
    const float og = vdot(nrm, bs.l);
    const bool through = og < 0.0f;
    if (!bs.delta && (through ? og >= 0.0f : og <= 0.0f)) return false;
    V3 f = bs.w;
    if (through) f = vscale(f, b.eta * b.eta);
    P.beta = vmul(P.beta, f);
    P.thr = vmul(P.thr, f);
    if (depth >= 3) {
        const float pr = rfmin(vmax3(P.thr), 0.95f);
        if (pcg_float(P.st, inc) >= pr) return false;
        P.beta = vscale(P.beta, 1.0f / pr);
        P.thr = vscale(P.thr, 1.0f / pr);
    }
    P.o = hp;
    P.d = vnorm(bs.l);
    return true;
}

// the launch's rays (cnt[0] closest hit, cnt[1] towards the camera) and, COUNT, its node records
// and shapes (cnt[2], cnt[3]), summed by wave
template <bool COUNT>
__device__ __forceinline__ void light_counts(unsigned long long *cnt, uint32_t n_ext, uint32_t n_shadow, const WalkCnt &wc)
{
    unsigned long long v[4] = {n_ext, n_shadow, wc.n, wc.t};
    for (int k = 0; k < (COUNT ? 4 : 2); k++) {
        for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
        if (lane_id() == 0 && v[k]) atomicAdd(&cnt[k], v[k]);
    }
}

// a light path's state in the queue: 64 bytes
__device__ __forceinline__ void light_put(float4 *q, const LightState &P)
{
    q[0] = make_float4(P.o.x, P.o.y, P.o.z, __uint_as_float(P.path));
    q[1] = make_float4(P.d.x, P.d.y, P.d.z, 0.0f);
    q[2] = make_float4(P.beta.x, P.beta.y, P.beta.z, P.thr.x);
    q[3] = make_float4(P.thr.y, P.thr.z, __uint_as_float((uint32_t)P.st), __uint_as_float((uint32_t)(P.st >> 32)));
}
__device__ __forceinline__ LightState light_get(const float4 *q)
{
    const float4 a = q[0], b = q[1], c = q[2], d = q[3];
    LightState P;
    P.o = v3(a.x, a.y, a.z);
    P.path = __float_as_uint(a.w);
    P.d = v3(b.x, b.y, b.z);
    P.beta = v3(c.x, c.y, c.z);
    P.thr = v3(c.w, d.x, d.y);
    P.st = (uint64_t)__float_as_uint(d.z) | ((uint64_t)__float_as_uint(d.w) << 32);
    return P;
}

// 1. light paths g0 .. g0+n-1 (path g: pass g / nlp, path g % nlp): their emission and first
// vertex; those that go on appended to the queue (a wave's together: one atomic add a wave)
template <int BLK, int STK, bool COUNT, bool PC = false>
__global__ __launch_bounds__(BLK) void k_light_first(JobDev J, uint32_t nlp, uint64_t g0, uint32_t n, uint32_t *spill,
                                                    size_t spill_stride, int spill_per, const float *tab, uint32_t *overflow,
                                                    float4 *queue, uint32_t *qcount, unsigned long long *cnt)
{
    __shared__ uint32_t stk[STK][BLK];
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    const WalkStack S{&stk[0][threadIdx.x], BLK, spill + i, spill_stride, STK + spill_per};
    const LightCtx L = light_ctx(J, nlp);
    uint32_t n_ext = 0, n_shadow = 0;
    WalkCnt wc;
    bool on = false;
    LightState P;
    if (i < n) {
        const uint64_t g = g0 + i;
        const uint32_t pass = (uint32_t)(g / nlp);
        P.path = (uint32_t)(g % nlp);
        const uint64_t inc = ((uint64_t)P.path << 1u) | 1u;
        P.st = pcg_seed(P.path, 0x80000000u + J.s0 + pass, L.seed);
        // emit: a point of the lamp, a direction about its normal
        const float e1 = pcg_float(P.st, inc), e2 = pcg_float(P.st, inc);
        P.o = vadd(v3of(L.lamp.q), vadd(vscale(v3of(L.lamp.u), e1), vscale(v3of(L.lamp.v), e2)));
        const float c1 = pcg_float(P.st, inc), c2 = pcg_float(P.st, inc);
        P.d = cosine_hemisphere(v3of(L.lamp.n), c1, c2);
        P.beta = vscale(L.le, L.lamp.v.w * PI_F);
        P.thr = v3(1.0f, 1.0f, 1.0f);
        if (J.max_depth > 0)
            on = light_vertex<STK, COUNT, PC>(J, L, 0, P, S, tab, overflow, wc, n_ext, n_shadow, nullptr) &&
                 J.max_depth > 1;
    }
    const uint64_t b = __ballot(on);
    uint32_t at = 0;
    if (lane_id() == 0 && b) at = atomicAdd(qcount, (uint32_t)__popcll(b));
    at = __shfl(at, 0, 64);
    if (on) light_put(queue + 4 * (size_t)(at + __popcll(b & lanemask_lt())), P);
    light_counts<COUNT>(cnt, n_ext, n_shadow, wc);
}

// 2. the queue's light paths from their second vertex on, one a lane (PC: their deposits summed
// in the work-group first; work-groups past the queue's end leave at once)
template <int BLK, int STK, bool COUNT, bool PC = false>
__global__ __launch_bounds__(BLK) void k_light_rest(JobDev J, uint32_t nlp, uint32_t *spill, size_t spill_stride, int spill_per,
                                                   const float *tab, uint32_t *overflow, const float4 *queue,
                                                   const uint32_t *qcount, unsigned long long *cnt)
{
    __shared__ uint32_t stk[STK][BLK];
    const size_t gtid = (size_t)blockIdx.x * BLK + threadIdx.x;
    const WalkStack S{&stk[0][threadIdx.x], BLK, spill + gtid, spill_stride, STK + spill_per};
    const LightCtx L = light_ctx(J, nlp);
    const uint32_t nq = *qcount;
    if ((size_t)blockIdx.x * BLK >= nq) return;   // the whole work-group: none of its lanes has a path
    uint32_t n_ext = 0, n_shadow = 0;
    WalkCnt wc;
    PcLds *pl = nullptr;
    if constexpr (PC) {
        __shared__ PcLds sums;
        pl = &sums;
        pc_lds_init(sums);
        __syncthreads();
    }
    for (size_t k = gtid; k < nq; k += (size_t)gridDim.x * BLK) {
        LightState P = light_get(queue + 4 * k);
        for (int depth = 1; depth < J.max_depth; depth++)
            if (!light_vertex<STK, COUNT, PC>(J, L, depth, P, S, tab, overflow, wc, n_ext, n_shadow, pl))
                break;
    }
    if constexpr (PC) {
        __syncthreads();
        pc_lds_flush(J, *pl);
    }
    light_counts<COUNT>(cnt, n_ext, n_shadow, wc);
}

__global__ void k_light_reset(uint32_t *qcount) { *qcount = 0u; }

template <int BLK, int STK, bool COUNT>
void L_light(hipStream_t s, const JobDev &J, uint32_t nlp, uint64_t g0, uint32_t n, uint32_t *spill, int spill_per,
             const float *tab, uint32_t *overflow, float4 *queue, uint32_t *qcount, unsigned long long *cnt)
{
    const unsigned blocks = (n + BLK - 1) / BLK;
    hipLaunchKernelGGL(k_light_reset, dim3(1), dim3(1), 0, s, qcount);
    hipLaunchKernelGGL((k_light_first<BLK, STK, COUNT>), dim3(blocks), dim3(BLK), 0, s, J, nlp, g0, n, spill,
                       (size_t)blocks * BLK, spill_per, tab, overflow, queue, qcount, cnt);
    hipLaunchKernelGGL((k_light_rest<BLK, STK, COUNT>), dim3(blocks), dim3(BLK), 0, s, J, nlp, spill, (size_t)blocks * BLK,
                       spill_per, tab, overflow, queue, qcount, cnt);
}

template <int BLK, int STK, bool TL>
void L_guide(dim3 g, hipStream_t s, const JobDev &J, float4 *guide, uint32_t *spill, size_t stride, int spill_per,
             const float *tab, uint32_t *overflow)
{
    hipLaunchKernelGGL((k_guide<BLK, STK, TL>), g, dim3(BLK), 0, s, J, guide, spill, stride, spill_per, tab, overflow);
}

template <int BLK>
void L_regen(dim3 g, hipStream_t s, const EngArgs &A)
{
    hipLaunchKernelGGL(k_regen<BLK>, g, dim3(BLK), 0, s, A);
}
template <int BLK, bool MULTI>
void L_shade(dim3 g, hipStream_t s, const EngArgs &A)
{
    hipLaunchKernelGGL((k_shade<BLK, MULTI>), g, dim3(BLK), 0, s, A);
}
template <int BLK, int STK, bool COUNT, bool TL>
void L_extend(dim3 g, hipStream_t s, const EngArgs &A)
{
    hipLaunchKernelGGL((k_extend<BLK, STK, COUNT, TL>), g, dim3(BLK), 0, s, A);
}
template <int BLK, int STK, bool COUNT>
void L_extend_rf(dim3 g, hipStream_t s, const EngArgs &A)
{
    hipLaunchKernelGGL((k_extend_rf<BLK, STK, COUNT>), g, dim3(BLK), 0, s, A);
}
template <int BLK, int STK, bool COUNT>
void L_shadow_rf(dim3 g, hipStream_t s, const EngArgs &A)
{
    hipLaunchKernelGGL((k_shadow_rf<BLK, STK, COUNT>), g, dim3(BLK), 0, s, A);
}
template <int BLK, int STK, bool COUNT, bool TL, bool MULTI>
void L_shadow(dim3 g, hipStream_t s, const EngArgs &A)
{
    hipLaunchKernelGGL((k_shadow<BLK, STK, COUNT, TL, MULTI>), g, dim3(BLK), 0, s, A);
}

template <int BLK, int STK, int K>
void fill_walks(LrEngTable &t)
{
    const int b = lr_blk_index(BLK);
    t.extend[b][K][0] = L_extend<BLK, STK, false, false>;
    t.extend[b][K][1] = L_extend<BLK, STK, false, true>;
    t.extend_count[b][K][0] = L_extend<BLK, STK, true, false>;
    t.extend_count[b][K][1] = L_extend<BLK, STK, true, true>;
    t.extend_rf[b][K][0] = L_extend_rf<BLK, STK, false>;
    t.extend_rf[b][K][1] = L_extend_rf<BLK, STK, true>;
    t.shadow_rf[b][K][0] = L_shadow_rf<BLK, STK, false>;
    t.shadow_rf[b][K][1] = L_shadow_rf<BLK, STK, true>;
    t.shadow[b][K][0][0] = L_shadow<BLK, STK, false, false, false>;
    t.shadow[b][K][1][0] = L_shadow<BLK, STK, false, true, false>;
    t.shadow[b][K][0][1] = L_shadow<BLK, STK, false, false, true>;
    t.shadow[b][K][1][1] = L_shadow<BLK, STK, false, true, true>;
    t.shadow_count[b][K][0][0] = L_shadow<BLK, STK, true, false, false>;
    t.shadow_count[b][K][1][0] = L_shadow<BLK, STK, true, true, false>;
    t.shadow_count[b][K][0][1] = L_shadow<BLK, STK, true, false, true>;
    t.shadow_count[b][K][1][1] = L_shadow<BLK, STK, true, true, true>;
}

template <int BLK>
void fill_table(LrEngTable &t)
{
    const int b = lr_blk_index(BLK);
    if (BLK == 256) {
        t.guide[0] = L_guide<256, 16, false>;
        t.guide[1] = L_guide<256, 16, true>;
        t.light[0] = L_light<256, 16, false>;
        t.light[1] = L_light<256, 16, true>;
    }
    t.regen[b] = L_regen<BLK>;
    t.shade[b][0] = L_shade<BLK, false>;
    t.shade[b][1] = L_shade<BLK, true>;
    fill_walks<BLK, 16, 0>(t);
    fill_walks<BLK, 24, 1>(t);
    fill_walks<BLK, 32, 2>(t);
}

}  // namespace
