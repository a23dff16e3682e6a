"""The oracle's photon cache (lro_pc_*, lro_photons_trace): the viewer's caustics between frames
(deposit / pc_claim, trace.wgsl:1359-1404; trace_decay.wgsl; trace_resolve.wgsl:43-80), pinned:

- closed form: photons spread uniformly over a floor, power 1 each, give a radiance of
  albedo / pi x (photons / paths) / area where the camera sees it: within 1 % when every finest
  cell holds 16 photons or more; the adaptive choice of cell size (the finest holding 16) leans
  high where they are fewer, by a few per cent;
- every photon lands in one cell of each of the three sizes, whatever the order of deposits;
- the fade: sums rounded, photons floored, a cell given back once its photons are gone;
- the light paths' deposits (lro_photons_trace) against the light tracer (lro_light_trace) on the
  caustics scene: the two estimate the same radiance on the rough surfaces the camera sees, the
  cache blurred over its cells.
"""
import ctypes

import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro


def tri(a, b, c, m=0):
    t = np.zeros(12, np.float32)
    t[0:3] = a
    t[4:7] = np.subtract(b, a)
    t[8:11] = np.subtract(c, a)
    t[[3, 7, 11]] = np.array([0xffffffff, 0, m << 26], np.uint32).view(np.float32)
    return t


def floor_scene(W, H):
    """a diffuse floor (albedo 0.5) at y = 0 over [-1, 2]^2, seen from above"""
    mats = (lr.Material * 1)()
    mats[0].type, mats[0].roughness, mats[0].ior = lr.MAT_DIFFUSE, 1.0, 1.5
    for k in range(3):
        mats[0].albedo[k] = 0.5
    tris = np.stack([tri((-1, 0, -1), (2, 0, -1), (-1, 0, 2)), tri((2, 0, 2), (-1, 0, 2), (2, 0, -1))])
    cam = lr.camera_make((0.5, 2.0, 0.5001), (0.5, 0.0, 0.5), (0, 0, 1), 20.0, W / H)
    return {"tris": tris, "prims": np.zeros((0, 24), np.float32), "mats": mats, "light": -1, "cam": cam, "eps": 1e-4,
            "margin": 1e-3}


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    xz = rng.random((n, 2))
    p = np.stack([xz[:, 0], np.zeros(n), xz[:, 1]], 1)
    return p, np.tile([0.0, 1.0, 0.0], (n, 1)), np.ones((n, 3))


@pytest.mark.parametrize("n,tol", [(2_000_000, 0.01), (200_000, 0.05)], ids=["dense", "sparse"])
def test_uniform_photons_closed_form(n, tol):
    W = H = 48
    sc = floor_scene(W, H)
    t = lro.pc_table()
    lro.pc_deposit_many(t, 0.01, *uniform(n, 1))
    out = lro.pc_resolve(sc, t, 0.01, float(n), W, H, lr.brdf_tables())
    assert np.all(out[..., 3] == 1)
    expect = 0.5 / np.pi   # albedo / pi x irradiance (n photons of power 1 over area 1, n paths)
    inner = out[6:-6, 6:-6, :3]
    assert abs(inner.mean() / expect - 1) < tol
    assert np.allclose(inner.mean((0, 1)), inner.mean(), rtol=1e-6)   # grey in, grey out


def test_every_photon_in_three_cells_in_any_order():
    p, nn, c = uniform(50_000, 2)
    c = c * np.linspace(0.5, 2.0, len(c))[:, None]
    a, b = lro.pc_table(), lro.pc_table()
    lro.pc_deposit_many(a, 0.02, p, nn, c)
    order = np.random.default_rng(3).permutation(len(p))
    lro.pc_deposit_many(b, 0.02, p[order], nn[order], c[order])
    ca, cb = lro.pc_cells(a), lro.pc_cells(b)
    assert ca == cb
    tot = np.array(list(ca.values()), np.int64).sum(0)
    assert tot[3] == 3 * 16 * len(p)
    per = np.rint(c.astype(np.float32) * np.float32(256.0)).astype(np.int64).sum(0)
    assert np.array_equal(tot[:3], 3 * per)


def test_fade_rounds_and_gives_cells_back():
    t = lro.pc_table()
    lro.pc_deposit(t, 0.05, [0.3, 0.0, 0.3], [0, 1, 0], [1.0, 0.5, 0.25])   # one photon: three cells
    lro.pc_deposit(t, 0.05, [0.7, 0.0, 0.7], [0, 1, 0], [3.0, 3.0, 3.0])
    before = lro.pc_cells(t)
    lro.pc_fade(t, 0.5)
    after = lro.pc_cells(t)
    assert after.keys() == before.keys()
    for g, (r, gg, b, n) in before.items():
        assert after[g] == (int(np.rint(r * 0.5)), int(np.rint(gg * 0.5)), int(np.rint(b * 0.5)), n // 2)
    for _ in range(5):   # 8 -> 4 -> 2 -> 1 -> 0 photon sixteenths: then the slots are given back
        lro.pc_fade(t, 0.5)
    assert len(lro.pc_cells(t)) == 0 and not t.any()


def test_cache_agrees_with_the_light_tracer():
    W = H = 64
    tab = lr.brdf_tables()
    sc = lr.caustic_scene(W, H)
    t = lro.pc_table()
    npaths = 2_000_000
    lro.photons_trace(sc, t, 0.01, npaths, 0, 3, 8, tab)
    pc = lro.pc_resolve(sc, t, 0.01, float(npaths), W, H, tab)
    npass = 1024
    light = lro.light_trace(sc, W, H, W * H, 0, npass, seed=5, tab=tab) / 2.0 ** 28 / npass
    seen = pc[..., 3] > 0
    assert seen.mean() > 0.5

    def blocks(x):
        return (x * seen[..., None]).reshape(H // 8, 8, W // 8, 8, 3).mean((1, 3, 4))

    a, b = blocks(pc[..., :3]), blocks(light)
    ok = b > 1e-4
    ratio = a[ok] / b[ok]
    # the cache spreads the glass ball's sharp caustic over its cells (and into pixels that see
    # the floor beside it): the bulk of the blocks agree, the median closely
    assert abs(np.median(ratio) - 1) < 0.08 and np.mean(np.abs(ratio - 1) < 0.15) > 0.7
