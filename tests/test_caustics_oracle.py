"""The oracle's light tracing for caustics (lro_light_trace) and the path tracer's skip_caustics,
on the caustics scene (a glass ball and a mirror ball in the Cornell room), pinned three ways:

- the split is the path tracer: plain = skip_caustics + light, block by block over 16 seeds (the
  plain path tracer is pinned by closed forms and against the reference elsewhere); the same
  test with the light part scaled by 2/3 or 3/2 (a lost pi, eta^2 or pixel area is a larger
  factor) must fail, which shows it can;
- the light image against the reference's own light_trace_caustics (oracle/_ref, its double
  build), block by block;
- the path tracer with skip_caustics against the reference's trace_image with skip_caustics.
Different random numbers on either side: z-scores of block means over independent seeds.
"""
import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro

caustic_scene = lr.caustic_scene

W = H = 32
K, SPP = 16, 256


def blocks(x):
    return np.asarray(x, np.float64).reshape(H // 8, 8, W // 8, 8, 3).mean((1, 3))


def zscores(a, b):
    """block z-scores and the whole image's z of two stacks of K independent estimates"""
    a, b = np.asarray(a), np.asarray(b)
    z = (a.mean(0) - b.mean(0)) / np.sqrt((a.var(0, ddof=1) + b.var(0, ddof=1)) / len(a) + 1e-30)
    ta, tb = a.mean((1, 2, 3)), b.mean((1, 2, 3))
    zt = (ta.mean() - tb.mean()) / np.sqrt((ta.var(ddof=1) + tb.var(ddof=1)) / len(a) + 1e-30)
    return z, zt


@pytest.fixture(scope="module")
def setup():
    return caustic_scene(W, H), lr.brdf_tables(), lro.blue_noise()


def light(sc, tab, mask, seed):
    return lro.light_trace(sc, W, H, W * H, 0, SPP, seed=seed, tab=tab, mask=mask) / 2.0 ** 28 / SPP


def test_skip_plus_light_is_the_path_tracer(setup):
    sc, tab, mask = setup
    plain = [blocks(lro.pt_image(sc, W, H, 0, SPP, seed=100 + k, tab=tab, mask=mask) / SPP) for k in range(K)]
    skip = [blocks(lro.pt_image(dict(sc, skip_caustics=1), W, H, 0, SPP, seed=300 + k, tab=tab, mask=mask) / SPP)
            for k in range(K)]
    lt = [blocks(light(sc, tab, mask, 300 + k)) for k in range(K)]
    share = np.mean(lt) / np.mean(plain)
    assert 0.005 < share < 0.1                      # the caustics are there, and a small part
    z, zt = zscores(plain, np.add(skip, lt))
    assert np.abs(z).max() < 5 and abs(zt) < 4
    # the light part was left out by the path tracer: without it the means differ
    _, zt0 = zscores(plain, skip)
    assert zt0 > 8
    for f in (2.0 / 3.0, 1.5):
        zf, ztf = zscores(plain, np.add(skip, np.multiply(lt, f)))
        assert np.abs(zf).max() > 5 or abs(ztf) > 4


def test_light_tracer_matches_reference(setup):
    sc, tab, mask = setup
    if lro.ref(True) is None:
        pytest.skip("oracle/_ref not built (the reference's sources are not here)")
    ours = [blocks(light(sc, tab, mask, 500 + k)) for k in range(K)]
    theirs = [blocks(lro.ref_render_lt(sc, W, H, SPP, seed=700 + k, trace=False, photons=1.0)) for k in range(K)]
    assert np.mean(theirs) > 0
    z, zt = zscores(ours, theirs)
    assert np.abs(z).max() < 5 and np.mean(np.abs(z) > 3) < 0.1 and abs(zt) < 4


def test_skip_caustics_matches_reference(setup):
    sc, tab, mask = setup
    if lro.ref(True) is None:
        pytest.skip("oracle/_ref not built (the reference's sources are not here)")
    ours = [blocks(lro.pt_image(dict(sc, skip_caustics=1), W, H, 0, SPP, seed=1100 + k, tab=tab, mask=mask) / SPP)
            for k in range(K)]
    theirs = [blocks(lro.ref_render_lt(sc, W, H, SPP, seed=900 + k, skip=True)) for k in range(K)]
    z, zt = zscores(ours, theirs)
    assert np.abs(z).max() < 5 and abs(zt) < 4


def test_light_tracer_is_order_free_and_split_over_passes(setup):
    # integer splats: the same sums in any thread order; passes s0.. of one call = calls of one
    # pass each
    sc, tab, mask = setup
    a = lro.light_trace(sc, W, H, 500, 3, 4, seed=9, tab=tab, mask=mask)
    b = lro.light_trace(sc, W, H, 500, 3, 4, seed=9, tab=tab, mask=mask)
    c = sum(lro.light_trace(sc, W, H, 500, 3 + k, 1, seed=9, tab=tab, mask=mask) for k in range(4))
    assert a.sum() > 0 and np.array_equal(a, b) and np.array_equal(a, c)
    # no lamp: nothing
    assert lro.light_trace(dict(sc, light=-1), W, H, 500, 0, 1, seed=9, tab=tab, mask=mask).sum() == 0
