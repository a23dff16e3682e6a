"""GPU tests of a job's target error (LrhJob.target_error, the serving workload's "target error of
1e-3"): its first 4 passes (1 sample a pixel each) decide the pass count K at which the image
mean's relative standard error reaches the target, K = ceil(4 (e4 / target)^2) within [4, spp].

- K equals the one recomputed here from the 4 passes rendered as jobs of their own (their image
  means; the device sums each sample's luminance in fixed point, so e4 agrees to ~1e-6 and K to
  within one pass where the ceiling falls close), and the reported error is e4 sqrt(4 / K).
- The job's image is exactly a K-spp job's (the same samples, the same sums), batched and plain.
- A loose target stops at 4 passes, an unreachable one runs all spp, no target runs all spp.
"""
import math
import sys

import numpy as np
import pytest

import little_renderer_amd as lr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    assert torch.cuda.is_available(), "these tests need the GPU"
    ctx = lr.Context(0)
    W, H = 160, 120
    sc = lr.cornell(W, H)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    eng = lr.Engine(ctx, slots_log2=18, max_tree_depth=ps.info()[0], image_bytes=256 << 20)
    yield ctx, ps, eng, sc["cam"], W, H
    ctx.close()


def run(eng, ps, cam, W, H, spp, s0=0, target=0.0, plain=False):
    j = lr.make_job(ps, cam, W, H, spp, s0=s0, target_error=target)
    i = eng.plain(j) if plain else eng.submit(j)
    if not plain:
        eng.drain()
    info = eng.job(i)
    t = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    eng.image(i, t)
    eng.release(i)
    return t.cpu().numpy(), info


def lum(img):
    return 0.2126 * img[..., 0].astype(np.float64) + 0.7152 * img[..., 1] + 0.0722 * img[..., 2]


def test_target_error_picks_the_pass_count_and_its_image(setup):
    ctx, ps, eng, cam, W, H = setup
    spp = 64
    mu = np.array([lum(run(eng, ps, cam, W, H, 1, s0=k)[0]).mean() for k in range(4)])
    e4 = math.sqrt(mu.var(ddof=1) / 4) / mu.mean()
    target = e4 / 2.3           # K = ceil(4 * 5.29) = 22 (away from a ceiling's edge)
    want = min(max(math.ceil(4 * (e4 / target) ** 2), 4), spp)
    img, info = run(eng, ps, cam, W, H, spp, target=target)
    say(f"  e4 {e4:.3e}, target {target:.3e}: K {info.spp_done} (recomputed {want}), reported error {info.rel_error:.3e}")
    assert abs(info.spp_done - want) <= 1
    assert abs(info.rel_error - e4 * math.sqrt(4 / info.spp_done)) < 1e-4 * e4
    assert info.rel_error <= target * 1.0001
    assert np.all(img[..., 3] == info.spp_done)
    ref, rinfo = run(eng, ps, cam, W, H, info.spp_done)
    assert rinfo.spp_done == info.spp_done
    assert np.array_equal(img, ref), "not the image of a job of that many samples"
    pimg, pinfo = run(eng, ps, cam, W, H, spp, target=target, plain=True)
    assert pinfo.spp_done == info.spp_done and np.array_equal(pimg, img)


def test_loose_unreachable_and_no_target(setup):
    ctx, ps, eng, cam, W, H = setup
    _, a = run(eng, ps, cam, W, H, 16, target=0.5)
    _, b = run(eng, ps, cam, W, H, 16, target=1e-9)
    _, c = run(eng, ps, cam, W, H, 16)
    say(f"  loose: {a.spp_done}, unreachable: {b.spp_done} ({b.rel_error:.2e}), none: {c.spp_done}")
    assert a.spp_done == 4 and b.spp_done == 16 and c.spp_done == 16 and c.rel_error == 0.0
    # a job of 4 samples or fewer takes them all, the target aside
    _, d = run(eng, ps, cam, W, H, 4, target=0.5)
    assert d.spp_done == 4
