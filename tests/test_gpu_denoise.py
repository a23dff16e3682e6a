"""GPU parity of the denoiser (lrh_denoise; jobs submitted with denoise = 1: their guide and moments).

- The guide (the first diffuse surface through each pixel's centre, through mirrors and glass)
  equals the oracle's (lro_guide) on the Cornell box and on a GGX scene with glass: the same walk
  and the same float operations, so to the bit but where the device's sqrt or its walk's order
  ties differently (a handful of pixels at shapes' seams).
- The moments of the samples' irradiance luminance equal the oracle's (lro_pt_image_moments on the
  device's guide) to its fixed point (2^-28 and 2^-20 units) where the samples agree.
- The filter on the same inputs equals the oracle's (lro_denoise): their exp and sqrt may round
  differently in the last bit, which moves a packed half by one step now and then.
- lrh_engine_denoise equals lrh_denoise on the job's image, moments and guide; a denoised job's
  image, guide and moments are the same batched and plain.
- It denoises: at 4 spp the filtered Cornell box is far closer to the converged image than the
  samples alone are, below the ceiling. (On the ceiling it is not: a pixel whose centre sees the
  ceiling and whose area is partly the lamp, hung just under the ceiling at nearly its depth and
  with its normal, is averaged with the ceiling around it, and the passes carry the lamp's light
  out across the ceiling. That is the reference's filter as it is: its guide and weights are the
  same there.)
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


def dev(shape):
    t = torch.zeros(shape, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    return t


def run_job(eng, ps, cam, W, H, spp, plain=False, seed=0):
    """a denoised job: (image (H, W, 4), guide (H, W, 8), moments (H, W, 2), denoised (H, W, 4)) on the device"""
    j = lr.make_job(ps, cam, W, H, spp, seed=seed, denoise=True)
    i = eng.plain(j) if plain else eng.submit(j)
    if not plain:
        eng.drain()
    img, g, m, den = dev((H, W, 4)), dev((H, W, 8)), dev((H, W, 2)), dev((H, W, 4))
    eng.image(i, img)
    eng.guide(i, g, m)
    eng.denoise(i, den)
    eng.release(i)
    return img, g, m, den


def scenes(W, H):
    cb = lr.cornell(W, H)
    gg = lr.batch_scene(3000, 7)
    gg["cam"] = lr.batch_camera(2, W, H)
    return [("cornell", cb), ("ggx-glass", gg)]


@pytest.mark.parametrize("which", [0, 1], ids=["cornell", "ggx-glass"])
def test_guide_moments_and_filter_match_the_oracle(ctx, which):
    W = H = 48
    spp = 8
    name, sc = scenes(W, H)[which]
    say(f"denoiser parity: {name} {W}x{H} spp {spp}")
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=ps.info()[0], image_bytes=64 << 20)
    img, g, m, den = (t.cpu().numpy() for t in run_job(eng, ps, sc["cam"], W, H, spp))
    tab = lr.brdf_tables()
    go = lro.guide(sc, W, H, tab)
    diff = np.abs(g - go).max(-1)
    say(f"  guide: {np.mean(diff == 0):.4f} of pixels equal to the bit, {np.mean(diff > 1e-4)} beyond 1e-4")
    assert np.mean(diff > 1e-4) < 0.01
    assert np.array_equal(g[..., 7], go[..., 7]) or np.mean(g[..., 7] != go[..., 7]) < 0.01
    so, mo = lro.pt_image_moments(sc, W, H, 0, spp, g, max_depth=8, tab=tab, mask=lro.blue_noise())
    rel = np.abs(m - mo) / np.maximum(np.abs(mo), 1e-3)
    say(f"  moments: median rel {np.median(rel):.2e}, >1e-3 in {np.mean(rel > 1e-3):.4f}")
    assert np.median(rel) < 1e-5 and np.mean(rel > 1e-3) < 0.02
    # the filter on the device's own inputs, both ways
    out = lro.denoise(img, m, g, passes=5)
    rel = np.abs(den[..., :3] - out[..., :3]) / np.maximum(np.abs(out[..., :3]), 1e-3)
    say(f"  filter: median rel {np.median(rel):.2e}, 99.9th pct {np.percentile(rel, 99.9):.2e}, max {rel.max():.2e}")
    assert np.median(rel) < 1e-5 and np.percentile(rel, 99.9) < 2e-3 and rel.max() < 2e-2
    assert np.array_equal(den[..., 3], out[..., 3])
    eng.close()
    ps.close()


def test_engine_denoise_is_lrh_denoise_and_batched_equals_plain(ctx):
    W, H = 64, 40
    sc = lr.cornell(W, H)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=ps.info()[0], image_bytes=64 << 20)
    img, g, m, den = run_job(eng, ps, sc["cam"], W, H, 4)
    out = dev((H, W, 4))
    ctx.denoise(img, m, g, W, H, out)
    ctx.sync()
    assert torch.equal(out, den)
    img2, g2, m2, den2 = run_job(eng, ps, sc["cam"], W, H, 4, plain=True)
    for a, b in ((img, img2), (g, g2), (m, m2), (den, den2)):
        assert torch.equal(a, b)
    eng.close()
    ps.close()


def test_it_denoises(ctx):
    W = H = 96
    sc = lr.cornell(W, H)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    eng = lr.Engine(ctx, slots_log2=18, max_tree_depth=ps.info()[0], image_bytes=256 << 20)
    img, g, m, den = (t.cpu().numpy() for t in run_job(eng, ps, sc["cam"], W, H, 4, seed=1))
    ref = dev((H, W, 4))
    i = eng.submit(lr.make_job(ps, sc["cam"], W, H, 1024, seed=2))
    eng.drain()
    eng.image(i, ref)
    eng.release(i)
    ref = ref.cpu().numpy()
    conv = ref[..., :3] / ref[..., 3:4]
    raw = img[..., :3] / img[..., 3:4]
    # the image below the ceiling (its top third and the lamp left out)
    keep = np.zeros((H, W), bool)
    keep[H // 3:, :] = True
    keep &= g[..., 7] == 0
    e_raw = np.sqrt(np.mean((raw - conv)[keep] ** 2))
    e_den = np.sqrt(np.mean((den[..., :3] - conv)[keep] ** 2))
    say(f"  RMSE against 1024 spp below the ceiling: the 4 samples {e_raw:.4f}, denoised {e_den:.4f}")
    assert e_den < 0.5 * e_raw
    # the filter keeps the image's energy (within a few percent) there and over the whole image
    assert abs(den[..., :3][keep].mean() - conv[keep].mean()) < 0.03 * conv[keep].mean()
    assert abs(den[..., :3].mean() - conv.mean()) < 0.03 * conv.mean()
    eng.close()
    ps.close()
