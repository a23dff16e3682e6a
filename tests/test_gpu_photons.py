"""GPU photon cache (lrh_photons_*: gpu_trace_photons, trace_gpu.c:1964-1996; deposit / pc_claim,
trace.wgsl:1359-1404; trace_decay.wgsl; trace_resolve.wgsl:43-80) against the oracle
(oracle/lro_pt.c, pinned by tests/test_photons_oracle.py) on the caustics scene:

- the table: the device's light paths are the oracle's (the same PCG32 streams), so every cell
  holds the same integer sums and photons, found by its tag (which slot a cell took can differ
  where two cells meet: the device claims in the hardware's order);
- two traces with the same numbers hold twice the sums, to the bit (integer sums);
- the fade, and the resolve (each pixel's lookup) against the oracle's on the same cells;
- refusals: a scene of instances (no lamp), a fade factor outside [0, 1].
"""
import sys

import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CELL = 0.01


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "these tests need the GPU"
    c = lr.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scene(ctx):
    sc = lr.caustic_scene(64, 64)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    yield sc, ps
    ps.close()


def same_cells(a, b):
    """the share of cells in either table whose sums and photons are the same in both"""
    keys = set(a) | set(b)
    return sum(1 for k in keys if a.get(k) == b.get(k)) / max(len(keys), 1)


def test_table_equals_the_oracle(ctx, scene):
    sc, ps = scene
    n = 300_000
    say(f"photons: {n} light paths vs oracle")
    pc = lr.Photons(ctx, CELL)
    pc.trace(ps, n, s0=2, seed=7, max_depth=8)
    g = lro.pc_cells(pc.read())
    t = lro.pc_table()
    lro.photons_trace(sc, t, CELL, n, 2, 7, 8, lr.brdf_tables())
    o = lro.pc_cells(t)
    share = same_cells(g, o)
    tg = np.array(list(g.values()), np.float64).sum(0)
    to = np.array(list(o.values()), np.float64).sum(0)
    say(f"  cells {len(g)} / {len(o)}, the same in {share:.5f}; photons {tg[3] / 16:.0f} / {to[3] / 16:.0f}")
    assert len(o) > 1000 and share > 0.99
    assert np.allclose(tg, to, rtol=1e-4)
    assert abs(pc.paths() - n) < 1e-9
    pc.close()


def test_traces_add_and_the_fade_matches(ctx, scene):
    sc, ps = scene
    n = 100_000
    pc = lr.Photons(ctx, CELL)
    pc.trace(ps, n, seed=3)
    one = pc.read()
    pc.trace(ps, n, seed=3)
    two = pc.read()
    held = (one[:, 3] >> np.uint64(32)) != 0
    assert held.sum() > 1000
    assert np.array_equal(two[held, :3], 2 * one[held, :3])
    assert np.array_equal(two[:, 3] & np.uint64(0xffffffff), 2 * (one[:, 3] & np.uint64(0xffffffff)))
    pc.fade(0.375)
    dev = pc.read()
    lro.pc_fade(two, 0.375)
    assert np.array_equal(dev, two)
    assert abs(pc.paths() - 2 * n * 0.375) < 1e-6
    pc.clear()
    assert not pc.read().any() and pc.paths() == 0.0
    pc.close()


def test_resolve_equals_the_oracle(ctx, scene):
    sc, ps = scene
    W = H = 64
    n = 1_000_000
    say(f"photons: resolve {W}x{H} after {n} light paths vs oracle")
    pc = lr.Photons(ctx, CELL)
    pc.trace(ps, n, seed=11)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    pc.resolve(ps, sc["cam"], W, H, out)
    g = out.cpu().numpy()
    o = lro.pc_resolve(sc, pc.read(), CELL, float(n), W, H, lr.brdf_tables())
    pc.close()
    assert np.array_equal(g[..., 3] > 0, o[..., 3] > 0) and (o[..., 3] > 0).mean() > 0.5
    rel = np.abs(g[..., :3] - o[..., :3]) / np.maximum(np.abs(o[..., :3]), 1e-6)
    say(f"  caustics mean {o[..., :3].mean():.5f}; median rel {np.median(rel):.2e}, >1e-4 in {np.mean(rel > 1e-4):.4f}")
    assert o[..., :3].mean() > 0 and np.median(rel) < 1e-5 and np.mean(rel > 1e-4) < 0.01


def test_refusals(ctx):
    pc = lr.Photons(ctx, CELL)
    with pytest.raises(lr.LrhError):
        pc.fade(1.5)
    mt = lr.uv_sphere(8, 5)
    inst = np.zeros(1, lr.INSTANCE_DTYPE)
    inst[0]["xform"] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    mats = (lr.Material * 1)()
    mats[0].type, mats[0].roughness = lr.MAT_DIFFUSE, 1.0
    tl = lr.PtScene.instanced(ctx, mt, np.array([0, len(mt)], np.int32), mats, inst)
    with pytest.raises(lr.LrhError):
        pc.trace(tl, 1000)
    tl.close()
    pc.close()
