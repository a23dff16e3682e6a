"""GPU parity: the HIP builders and walk against the oracle, through the C ABI.

Bit-exact: triangle boxes, bounds, Morton codes and their sorted order, LBVH and PLOC node and
aux arrays (topology and boxes), and every hit (t, id, barycentrics) of the walk over the same
tree. Hit ids against the CPU's SAH tree (a different tree) agree except on ties.
"""
import sys

import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "these tests need the GPU"
    c = lr.Context(0)
    yield c
    c.close()


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    torch.cuda.synchronize()
    return t


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def u32(t):
    return host(t).view(np.uint32)


def gpu_build(ctx, tris, builder, node0=0, ref0=0, ref_add=0, extra=0):
    n = tris.shape[0]
    t = dev(tris)
    boxes = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    ctx.tri_boxes(t, boxes)
    nodes = torch.zeros((node0 + 2 * n - 1 + extra, 8), dtype=torch.int32, device="cuda")
    aux = torch.zeros(ref0 + n + extra, dtype=torch.int32, device="cuda")
    srt = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    st = ctx.build(boxes, n, nodes, aux, builder, node0=node0, ref0=ref0, ref_add=ref_add, sorted_out=srt)
    ctx.sync()
    return dict(t=t, boxes=boxes, nodes=nodes, aux=aux, srt=srt, st=st)


def soup(n, seed):
    return lr.gen_soup(n, seed)


def test_tri_boxes_and_bounds_bit_exact(ctx):
    tris = soup(100_000, 2)
    say("boxes n=100000")
    t = dev(tris)
    boxes = torch.empty((tris.shape[0], 8), dtype=torch.float32, device="cuda")
    ctx.tri_boxes(t, boxes)
    b = torch.empty(6, dtype=torch.float32, device="cuda")
    ctx.bounds(boxes, b)
    ctx.sync()
    want = lro.tri_boxes(tris)
    assert np.array_equal(u32(boxes), want.view(np.uint32))
    lo, hi = lro.union(want)
    assert np.array_equal(host(b), np.concatenate([lo, hi]))


SIZES = [1, 2, 3, 5, 17, 100, 1000, 1025, 4095, 4096, 4097, 10_000, 100_003]


@pytest.mark.parametrize("builder", [lr.LBVH, lr.PLOC], ids=["lbvh", "ploc"])
@pytest.mark.parametrize("n", SIZES)
def test_build_bit_exact(ctx, builder, n):
    say(f"build builder={builder} n={n}")
    tris = soup(n, 1000 + n)
    g = gpu_build(ctx, tris, builder)
    nodes, aux, sc, si, depth = lro.build_tris(tris, ploc_=builder == lr.PLOC)
    srt = u32(g["srt"])
    assert np.array_equal(srt[:n], sc), "Morton codes / sort order"
    assert np.array_equal(srt[n:], si), "sorted items"
    assert np.array_equal(u32(g["aux"]), aux)
    got = u32(g["nodes"])
    assert np.array_equal(got, nodes), f"nodes differ at {np.nonzero((got != nodes).any(1))[0][:8]}"
    assert g["st"].depth == depth
    ok, cost, d = lro.check_tree(got, n)
    assert ok and d == depth
    if builder == lr.PLOC:   # the live clusters over the rounds (the roofline's count)
        assert g["st"].clusters == lro.ploc(sc, si, lro.tri_boxes(tris), live=True)[4]
    else:
        assert g["st"].clusters == 0


@pytest.mark.parametrize("builder", [lr.LBVH, lr.PLOC], ids=["lbvh", "ploc"])
def test_build_equal_codes(ctx, builder):
    # a third of the triangles share one centroid (equal codes: the index tie rule of delta),
    # the rest are spread
    say(f"build equal codes builder={builder}")
    tris = soup(30_000, 77)
    tris[::3] = tris[0]
    g = gpu_build(ctx, tris, builder)
    nodes, aux, sc, si, depth = lro.build_tris(tris, ploc_=builder == lr.PLOC)
    assert np.array_equal(u32(g["srt"])[30_000:], si)
    assert np.array_equal(u32(g["nodes"]), nodes) and np.array_equal(u32(g["aux"]), aux)


@pytest.mark.parametrize("builder", [lr.LBVH, lr.PLOC], ids=["lbvh", "ploc"])
def test_build_offsets(ctx, builder):
    # into a scene's shared arrays: nodes from node0, leaves' aux from ref0, ids + ref_add
    say(f"build offsets builder={builder}")
    n, node0, ref0, ref_add = 5000, 123, 45, 1000
    tris = soup(n, 5)
    g = gpu_build(ctx, tris, builder, node0=node0, ref0=ref0, ref_add=ref_add, extra=7)
    boxes = lro.tri_boxes(tris)
    lo, hi = lro.union(boxes)
    sc, si = lro.sort(lro.morton(boxes, lo, lro.inv_extent(lo, hi)))
    if builder == lr.PLOC:
        nodes, aux, depth, _ = lro.ploc(sc, si, boxes, node0, ref0, ref_add)
    else:
        nodes, aux, depth = lro.lbvh(sc, si, boxes, node0, ref0, ref_add)
    got = u32(g["nodes"])
    assert not got[:node0].any() and not got[node0 + 2 * n - 1:].any()
    assert np.array_equal(got[node0:node0 + 2 * n - 1], nodes)
    ga = u32(g["aux"])
    assert not ga[:ref0].any() and np.array_equal(ga[ref0:ref0 + n], aux)


@pytest.mark.parametrize("n", [1_000_000, 4_000_000], ids=["1M", "4M"])
@pytest.mark.parametrize("builder", [lr.LBVH, lr.PLOC], ids=["lbvh", "ploc"])
def test_build_full_size_and_deterministic(ctx, builder, n):
    say(f"build n={n} builder={builder}")
    tris = soup(n, 2)
    a = gpu_build(ctx, tris, builder)
    b = gpu_build(ctx, tris, builder)
    na, nb = u32(a["nodes"]), u32(b["nodes"])
    assert np.array_equal(na, nb) and np.array_equal(u32(a["aux"]), u32(b["aux"]))
    nodes, aux, sc, si, depth = lro.build_tris(tris, ploc_=builder == lr.PLOC)
    assert np.array_equal(na, nodes) and np.array_equal(u32(a["aux"]), aux)
    assert a["st"].depth == depth
    say(f"  depth {depth} build {a['st'].ms:.3f} ms")


def test_lbvh_past_the_buffer_path(ctx):
    """34M primitives: (2n - 1) x 32 bytes of records past 2 GiB, where the climbs above the
    work-groups hand records over with 8-byte device-scope accesses instead of the 16-byte buffer
    ones. The sorted output is checked as such (order, a permutation, each code the oracle's
    Morton code of its item) and the tree against the oracle's LBVH over it, bit for bit."""
    n = 34_000_000
    assert (2 * n - 1) * 32 >= 2 ** 31
    say(f"LBVH n={n}")
    tris = soup(n, 41)
    g = gpu_build(ctx, tris, lr.LBVH)
    srt = u32(g["srt"])
    sc, si = srt[:n].copy(), srt[n:].copy()
    del srt
    boxes = host(g["boxes"])
    assert np.all(sc[1:] >= sc[:-1])
    tie = sc[1:] == sc[:-1]
    assert np.all(si[1:][tie] > si[:-1][tie])
    assert np.array_equal(np.bincount(si, minlength=n), np.ones(n, np.int64))
    lo, hi = lro.union(boxes)
    assert np.array_equal(lro.morton(boxes, lo, lro.inv_extent(lo, hi))[si], sc)
    nodes, aux, depth = lro.lbvh(sc, si, boxes)
    assert np.array_equal(u32(g["aux"]), aux)
    got = u32(g["nodes"])
    assert np.array_equal(got, nodes), f"nodes differ at {np.nonzero((got != nodes).any(1))[0][:8]}"
    assert g["st"].depth == depth
    say(f"  depth {depth} build {g['st'].ms:.3f} ms")


def aimed_rays(tris, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.3, 1.3, size=(n, 3)).astype(np.float32)
    tgt = rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32)
    k = rng.integers(0, tris.shape[0], size=n // 2)
    a, b = rng.uniform(0, 0.5, size=(2, n // 2, 1))
    tgt[: n // 2] = tris[k, 0:3] + a * tris[k, 4:7] + b * tris[k, 8:11]
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 1e-4
    r[:, 4:7] = d
    r[:, 7] = 1e30
    return r


@pytest.mark.parametrize("builder", [lr.LBVH, lr.PLOC], ids=["lbvh", "ploc"])
@pytest.mark.parametrize("any_hit", [False, True], ids=["closest", "any"])
def test_trace_bit_exact(ctx, builder, any_hit):
    n = 200_000
    say(f"trace builder={builder} any={any_hit}")
    tris = soup(n, 31)
    g = gpu_build(ctx, tris, builder)
    rays = aimed_rays(tris, 100_000, 4)
    r = dev(rays)
    hits = torch.zeros((rays.shape[0], 4), dtype=torch.float32, device="cuda")
    ctx.trace(g["nodes"], 0, g["aux"], g["t"], r, rays.shape[0], hits, g["st"].depth, any_hit=any_hit)
    ctx.sync()
    got = u32(hits)
    want = lro.trace_gpu_layout(u32(g["nodes"]), 0, u32(g["aux"]), tris, rays, any_hit=any_hit).view(np.uint32)
    assert (want[:, 1] != 0xFFFFFFFF).sum() > 20_000
    assert np.array_equal(got, want), f"{(got != want).any(1).sum()} rays differ"


def test_trace_ids_against_cpu_sah_tree(ctx):
    # a different tree (the CPU's SAH, engine/rt/bvh.c): the same closest triangle, except on ties
    n = 100_000
    say("trace vs SAH")
    tris = soup(n, 12)
    g = gpu_build(ctx, tris, lr.PLOC)
    rays = aimed_rays(tris, 20_000, 13)
    hits = torch.zeros((rays.shape[0], 4), dtype=torch.float32, device="cuda")
    ctx.trace(g["nodes"], 0, g["aux"], g["t"], dev(rays), rays.shape[0], hits, g["st"].depth)
    ctx.sync()
    got = host(hits)
    sah = lro.Sah(lro.tri_boxes(tris), 4, 16, wide=True).trace(tris, rays)
    sub = slice(0, 2000)
    _, second = lro.trace_brute(tris, rays[sub])
    ig, isah = lro.hit_ids(got), lro.hit_ids(sah)
    diff = np.nonzero(ig != isah)[0]
    # every difference is a tie (two candidates at the same t within 1e-5 relative)
    for k in diff:
        assert abs(got[k, 0] - sah[k, 0]) <= 1e-5 * abs(sah[k, 0]), k
    tie = np.abs(second - sah[sub, 0]) <= 1e-5 * np.abs(sah[sub, 0])
    assert np.all((ig[sub] == isah[sub]) | tie)
    same = (ig == isah) & (isah != 0xFFFFFFFF)
    assert np.array_equal(got[same].view(np.uint32), sah[same].view(np.uint32))
    say(f"  {len(diff)} tie(s) of {rays.shape[0]}")


def test_queue_trace_counts_from_device(ctx):
    n = 50_000
    say("trace queue")
    tris = soup(n, 3)
    g = gpu_build(ctx, tris, lr.LBVH)
    rays = aimed_rays(tris, 10_000, 5)
    cnt = dev(np.array([6000], np.int32))
    hits = torch.full((rays.shape[0], 4), -7.0, dtype=torch.float32, device="cuda")
    ctx.trace_queue(g["nodes"], 0, g["aux"], g["t"], dev(rays), rays.shape[0], cnt, hits, g["st"].depth)
    ctx.sync()
    got = u32(hits)
    want = lro.trace_gpu_layout(u32(g["nodes"]), 0, u32(g["aux"]), tris, rays[:6000]).view(np.uint32)
    assert np.array_equal(got[:6000], want)
    assert np.all(host(hits)[6000:] == -7.0)
