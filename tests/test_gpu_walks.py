"""GPU parity of the engine's own walks on one-level scenes (lrh_pt_scene_trace): the compressed
4-wide closest-hit walk and the any-hit walk, against the oracle's brute force over every shape.

The scene is the batched workload's (a triangle soup, a floor quad of two triangles and a quad
lamp primitive), built with both builders. Rays start around and inside the soup and aim at
shapes (triangles and the lamp), so most hit. The bar is the tie rule: the same hit or miss on
every ray, the same shape and t within 1e-5 relative except where two shapes are hit at the same
distance (then either may be reported, at that distance). Any-hit: the same occluded / free
answer on every ray.
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


def aimed_rays(sc, n, seed):
    """rays from points around and inside the unit cube toward shapes' centres (a tenth toward
    the lamp), tmin 1e-4, tmax 1e30"""
    rng = np.random.default_rng(seed)
    tris = np.asarray(sc["tris"], np.float32).reshape(-1, 12)
    o = rng.uniform(-0.3, 1.3, size=(n, 3)).astype(np.float32)
    pick = tris[rng.integers(0, len(tris), n)]
    tgt = pick[:, 0:3] + (pick[:, 4:7] + pick[:, 8:11]) / 3
    lamp = np.asarray(sc["prims"], np.float32).reshape(-1, 24)[0]
    k = rng.random(n) < 0.1
    tgt[k] = lamp[4:7] + 0.5 * lamp[8:11] + 0.5 * lamp[12:15]
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 1e-4, d, 1e30
    return rays


@pytest.mark.parametrize("builder", [lr.LBVH, lr.PLOC])
def test_one_level_walks_match_brute_force(ctx, builder):
    sc = lr.batch_scene(6000, 17)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], 0, sc["eps"], sc["margin"], builder=builder)
    rays = aimed_rays(sc, 20000, 5)
    say(f"builder {builder}: {len(sc['tris'])} triangles, {rays.shape[0]} rays")
    osc = {"prims": sc["prims"], "tris": sc["tris"], "mats": sc["mats"], "eps": sc["eps"], "light": 0}
    want = lro.scene_trace(osc, rays)
    got = ps.trace(rays)
    wi, gi = want[:, 1].view(np.uint32), got[:, 1].view(np.uint32)
    hit = wi != 0xffffffff
    assert hit.mean() > 0.6
    assert np.array_equal(hit, gi != 0xffffffff)
    rel = np.abs(got[hit, 0] - want[hit, 0]) / want[hit, 0]
    same = wi[hit] == gi[hit]
    ties = ~same
    say(f"  closest: {hit.sum()} hits, {ties.sum()} ties, lamp hits {(gi[hit] == 0).sum()}, "
        f"max rel t (same shape) {rel[same].max():.2e}")
    assert rel[same].max() < 1e-5
    assert rel[ties].max(initial=0.0) < 1e-5          # a tie: another shape at the same distance
    assert ties.mean() < 0.002
    assert (gi[hit] == 0).sum() > 100                  # the lamp primitive is found
    wa = lro.scene_trace(osc, rays, any_hit=True)
    ga = ps.trace(rays, any_hit=True)
    occ_w, occ_g = wa[:, 1].view(np.uint32) != 0xffffffff, ga[:, 1].view(np.uint32) != 0xffffffff
    say(f"  any hit: {occ_w.sum()} occluded")
    assert np.array_equal(occ_w, occ_g)
