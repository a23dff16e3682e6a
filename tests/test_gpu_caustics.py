"""GPU light tracing for caustics (LrhJob.photons: light_trace_caustics, tracer.c:511-665, with
the path tracer's skip_caustics, tracer.c:371-375) on the caustics scene (a glass ball and a
mirror ball in the Cornell room).

- Against the oracle on the same random numbers: the device's image is the oracle's path tracer
  with skip_caustics plus the oracle's light tracer (lro_light_trace, the same PCG32 streams and
  the same fixed-point splats), sample for sample; the light part is ~2% of the image and lands
  on a quarter of the pixels, so a wrong light tracer shows in a quarter of them.
- Batched = plain = lrh_pt_render_job to the bit with light tracing on (its splats are integer
  sums).
- Unbiased: over 16 seeds the device's plain image and its split (skip + light) agree block by
  block (8 x 8 pixels, the channels' mean: they are one test three times), and the split is less
  noisy. (The plain image's caustics are heavy-tailed: with few seeds a block's z strays; at 96
  seeds on the oracle every block's |z| < 3.)
- A scene with no mirror or glass: photons change nothing. Refusals: photons < 0, photons with a
  target error.
"""
import sys

import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

caustic_scene = lr.caustic_scene


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "these tests need the GPU"
    c = lr.Context(0)
    yield c
    c.close()


def images(eng, ids, jobs):
    out = []
    for i, j in zip(ids, jobs):
        t = torch.zeros((j.h, j.w, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        eng.image(i, t)
        out.append(t.cpu().numpy())
        eng.release(i)
    return out


def batched(eng, jobs):
    ids = [eng.submit(j) for j in jobs]
    assert all(i is not None for i in ids)
    eng.drain()
    return images(eng, ids, jobs)


def scene_on_device(ctx, sc):
    return lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])


def test_light_tracing_matches_the_oracle(ctx):
    W = H = 48
    spp, seed, photons = 8, 5, 1.0
    say(f"caustics: {W}x{H}x{spp}, {photons} light paths a pixel and pass, vs oracle")
    sc = caustic_scene(W, H)
    ps = scene_on_device(ctx, sc)
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=64, image_bytes=64 << 20)
    job = lr.make_job(ps, sc["cam"], W, H, spp, seed=seed, photons=photons)
    g = batched(eng, [job])[0]
    s1 = eng.stats()
    # again with the counting walks and the kernels timed: the same image
    eng.counting(True)
    eng.profiling(True)
    g2 = batched(eng, [job])[0]
    s2 = eng.stats()
    eng.close()
    ps.close()
    assert np.all(g[..., 3] == spp)
    assert np.array_equal(g.view(np.uint32), g2.view(np.uint32))
    nlp = int(photons * W * H)
    assert s1.light_launches == 1 and s1.light_paths == nlp * spp
    assert s1.light_rays >= s1.light_paths and s1.light_shadow_rays > 0 and s1.light_node_fetches == 0
    assert s2.light_rays - s1.light_rays == s1.light_rays and s2.light_shadow_rays - s1.light_shadow_rays == s1.light_shadow_rays
    assert s2.light_node_fetches > 0 and s2.light_shape_tests > 0 and s2.light_ms > 0
    say(f"  light paths {s1.light_paths}: rays {s1.light_rays} + {s1.light_shadow_rays} towards the camera, "
        f"{s2.light_node_fetches / s1.light_rays:.1f} nodes a ray (both kinds), {s2.light_ms:.3f} ms")
    tab, mask = lr.brdf_tables(), lro.blue_noise()
    pt = lro.pt_image(dict(sc, skip_caustics=1), W, H, 0, spp, seed=seed, tab=tab, mask=mask)
    light = lro.light_trace(sc, W, H, nlp, 0, spp, seed=seed, tab=tab, mask=mask) / 2.0 ** 28
    o = pt + light
    assert light.sum() > 0.01 * o.sum() and np.mean(light.sum(-1) > 0) > 0.2
    rel = np.abs(g[..., :3] - o) / np.maximum(np.abs(o), 1e-3)
    say(f"  light share {light.sum() / o.sum():.4f}; median rel {np.median(rel):.2e}, >1e-3 in {np.mean(rel > 1e-3):.4f}")
    assert np.median(rel) < 2e-5 and np.mean(rel > 1e-3) < 0.02
    assert abs(g[..., :3].mean() - o.mean()) < 1e-3 * o.mean()


def test_batched_equals_plain_with_light_tracing(ctx):
    sc = caustic_scene(64, 64)
    cb = lr.cornell(64, 64)
    ps = scene_on_device(ctx, sc)
    pc = scene_on_device(ctx, cb)
    jobs = [lr.make_job(ps, sc["cam"], 64, 64, 6, seed=1, photons=1.0),
            lr.make_job(ps, caustic_scene(40, 56)["cam"], 40, 56, 3, s0=7, seed=2, photons=0.37),
            lr.make_job(pc, cb["cam"], 64, 64, 4, seed=3, photons=2.0),
            lr.make_job(ps, sc["cam"], 64, 64, 5, seed=4)]
    say("caustics: batched vs plain, 4 jobs")
    eng = lr.Engine(ctx, slots_log2=17, max_tree_depth=64, image_bytes=64 << 20)
    a = batched(eng, jobs)
    b = images(eng, [eng.plain(j) for j in jobs], jobs)
    # no mirror or glass: photons change nothing
    c = images(eng, [eng.plain(lr.make_job(pc, cb["cam"], 64, 64, 4, seed=3))], [jobs[2]])[0]
    eng.close()
    # lrh_pt_render_job (the context's own engine): the same image
    d = torch.zeros((64, 64, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ps.render_job(jobs[0], d)
    ctx.sync()
    assert np.array_equal(d.cpu().numpy().view(np.uint32), a[0].view(np.uint32))
    ps.close()
    pc.close()
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.array_equal(a[2].view(np.uint32), c.view(np.uint32))


def test_split_is_unbiased_and_less_noisy(ctx):
    W = H = 64
    K, spp = 16, 256
    say(f"caustics: plain vs split, {K} seeds x {W}x{H}x{spp}")
    sc = caustic_scene(W, H)
    ps = scene_on_device(ctx, sc)
    eng = lr.Engine(ctx, slots_log2=18, max_tree_depth=64, image_bytes=256 << 20)
    plain = batched(eng, [lr.make_job(ps, sc["cam"], W, H, spp, seed=100 + k) for k in range(K)])
    split = batched(eng, [lr.make_job(ps, sc["cam"], W, H, spp, seed=200 + k, photons=1.0) for k in range(K)])
    eng.close()
    ps.close()

    def blocks(im):
        return (im[..., :3].astype(np.float64) / spp).reshape(H // 8, 8, W // 8, 8, 3).mean((1, 3, 4))

    P = np.array([blocks(x) for x in plain])
    S = np.array([blocks(x) for x in split])
    z = (P.mean(0) - S.mean(0)) / np.sqrt((P.var(0, ddof=1) + S.var(0, ddof=1)) / K)
    tp, ts = P.mean((1, 2)), S.mean((1, 2))
    zt = (tp.mean() - ts.mean()) / np.sqrt((tp.var(ddof=1) + ts.var(ddof=1)) / K)
    ratio = np.median(P.std(0, ddof=1) / np.maximum(S.std(0, ddof=1), 1e-12))
    say(f"  block |z| max {np.abs(z).max():.2f}, >3 in {np.mean(np.abs(z) > 3):.3f}; image z {zt:.2f}; noise ratio {ratio:.2f}")
    assert np.abs(z).max() < 4.5 and np.sum(np.abs(z) > 3) <= 3 and abs(zt) < 4
    assert ratio > 1.3


def test_photons_refused(ctx):
    sc = caustic_scene(16, 16)
    ps = scene_on_device(ctx, sc)
    eng = lr.Engine(ctx, slots_log2=14, max_tree_depth=64, image_bytes=16 << 20)
    with pytest.raises(lr.LrhError):
        eng.submit(lr.make_job(ps, sc["cam"], 16, 16, 4, photons=-1.0))
    with pytest.raises(lr.LrhError):
        eng.submit(lr.make_job(ps, sc["cam"], 16, 16, 4, photons=1.0, target_error=0.01))
    eng.close()
    ps.close()


def test_light_chunks_split_invariant(ctx):
    # 128 x 128, one light path a pixel and pass: 300 passes are 4.9M light paths, two launch
    # pairs (chunks of 2^22 paths: the second from path 2^22 on); 256 passes are exactly one chunk.
    # Every number depends only on (pass, path) and (pixel, sample), and the sums are integers:
    # passes 0..299 equal passes 0..255 plus passes 256..299 (to the float readout's rounding)
    W = H = 128
    sc = caustic_scene(W, H)
    ps = scene_on_device(ctx, sc)
    eng = lr.Engine(ctx, slots_log2=18, max_tree_depth=64, image_bytes=64 << 20)
    say("caustics: 300 passes (two light chunks) vs 256 + 44")
    a, b, c = batched(eng, [lr.make_job(ps, sc["cam"], W, H, 300, seed=8, photons=1.0),
                            lr.make_job(ps, sc["cam"], W, H, 256, seed=8, photons=1.0),
                            lr.make_job(ps, sc["cam"], W, H, 44, s0=256, seed=8, photons=1.0)])
    st = eng.stats()
    eng.close()
    ps.close()
    assert st.light_paths == W * H * (300 + 256 + 44) and st.light_launches == 3
    bc = b[..., :3].astype(np.float64) + c[..., :3]
    assert np.all(a[..., 3] == 300)
    assert np.allclose(a[..., :3], bc, rtol=1e-6, atol=1e-6)
