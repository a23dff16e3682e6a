"""GPU parity of the path tracer's several lights: glowing shapes sampled as lights (Scene.emitters,
engine/rt/tracer.c:120-229), glowing surfaces found by the paths, and a lamp with a sky in one
scene (tracer.c:467-478: all of them sampled at every rough vertex, a shadow ray each).

- Against the oracle sample for sample (the same Sobol pairs and PCG32 streams): a glowing ball
  and a glowing turned box over a floor, with MIS and without; a glowing ball under the sky; the
  batched workload's lamp-lit GGX scene under a sky too. The device's sin/cos/sqrt may round
  differently from glibc's in the last bit, so the median pixel agrees to ~1e-5 and a rare
  sample differs.
- The reference's closed form (tests/test_scene.c:118-153) on the device: the floor point below
  the ball reads 0.5 * 10 * (10 / 50)^2 = 0.2 within 0.004, with NEE and MIS and with NEE alone.
- Shade and shadow come in two forms (one shadow ray a vertex, or up to three): an image is the
  same to the bit whichever form the engine runs, so a lamp-lit job alone and beside a job with
  several lights give the same image, and batched equals plain for several lights too.
"""
import sys

import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro
from glow_scenes import glow_scene

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


def device_scene(ctx, parts, sky=None):
    mtris, first, mats, inst = parts
    ps = lr.PtScene.instanced(ctx, mtris, first, mats, inst, eps=2e-3)
    if sky is not None:
        ps.set_sky(sky)
    return ps


def render(eng, ps, cam, W, H, spp, seed=0, max_depth=8, mis=True, plain=False):
    j = lr.make_job(ps, cam, W, H, spp, seed=seed, max_depth=max_depth, mis=mis)
    i = eng.plain(j) if plain else eng.submit(j)
    if not plain:
        eng.drain()
    t = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    eng.image(i, t)
    eng.release(i)
    return t.cpu().numpy()


def compare(g, o, what):
    rel = np.abs(g[..., :3] - o) / np.maximum(np.abs(o), 1e-3)
    say(f"  {what}: median rel {np.median(rel):.2e}, >1e-3 in {np.mean(rel > 1e-3):.4f}, means {g[..., :3].mean():.5f} "
        f"{o.mean():.5f}")
    assert np.median(rel) < 2e-5
    assert np.mean(rel > 1e-3) < 0.02
    assert abs(g[..., :3].mean() - o.mean()) < 2e-3 * max(o.mean(), 1e-6)


def engine(ctx, depth):
    return lr.Engine(ctx, slots_log2=16, max_tree_depth=max(depth, 8), image_bytes=64 << 20)


@pytest.mark.parametrize("box", [False, True], ids=["ball", "box"])
@pytest.mark.parametrize("mis", [True, False], ids=["mis", "nee"])
def test_glowing_shape_matches_oracle(ctx, box, mis):
    W = H = 48
    spp = 8
    say(f"glowing {'box' if box else 'ball'} mis={mis} {W}x{H} spp {spp}")
    sc, parts = glow_scene(box=box, W=W, H=H)
    # a wider view: the floor around the point, the shape's glow and its shadowless side
    sc["cam"] = lr.camera_make((-200, 20, 0), (0, 20, 0), (0, 1, 0), 40.0, W / H)
    ps = device_scene(ctx, parts)
    eng = engine(ctx, ps.info()[0])
    g = render(eng, ps, sc["cam"], W, H, spp, mis=mis)
    o = lro.pt_image(sc, W, H, 0, spp, max_depth=8, mis=mis, tab=lr.brdf_tables(), mask=lro.blue_noise())
    compare(g, o, "engine vs oracle")
    eng.close()
    ps.close()


def test_glowing_ball_closed_form_on_device(ctx):
    n = 200
    sc, parts = glow_scene(box=False, W=n, H=n)
    ps = device_scene(ctx, parts)
    eng = engine(ctx, ps.info()[0])
    for mis in (True, False):
        g = render(eng, ps, sc["cam"], n, n, 1, mis=mis)
        v = g[..., 0].astype(np.float64)
        say(f"  glowing ball, mis={mis}: mean {v.mean():.5f}, sd {v.std():.4f} (0.5 * 10 * (10 / 50)^2 = 0.2)")
        assert abs(v.mean() - 0.2) < 0.004
    eng.close()
    ps.close()


def test_glowing_ball_under_the_sky_matches_oracle(ctx):
    W = H = 40
    spp = 8
    sky = lr.sky(128, 35.0, 60.0, sun_le=(50.0, 46.0, 40.0))
    sc, parts = glow_scene(box=False, W=W, H=H, sky=sky)
    sc["cam"] = lr.camera_make((-200, 40, 30), (0, 20, 0), (0, 1, 0), 50.0, W / H)
    ps = device_scene(ctx, parts, sky=sky)
    eng = engine(ctx, ps.info()[0])
    g = render(eng, ps, sc["cam"], W, H, spp)
    o = lro.pt_image(sc, W, H, 0, spp, max_depth=8, mis=True, tab=lr.brdf_tables(), mask=lro.blue_noise())
    compare(g, o, "glowing ball and sky: engine vs oracle")
    eng.close()
    ps.close()


def lamp_and_sky_scene():
    sc = lr.batch_scene(4000, 5)
    sc["sky"] = lro.Env(lr.sky(128, 30.0, 120.0, sun_le=(40.0, 36.0, 30.0)))
    return sc


def test_lamp_and_sky_matches_oracle(ctx):
    W = H = 48
    spp = 8
    sc = lamp_and_sky_scene()
    cam = lr.batch_camera(11, W, H)
    sc["cam"] = cam
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    ps.set_sky(sc["sky"].rgb)
    eng = engine(ctx, ps.info()[0])
    g = render(eng, ps, cam, W, H, spp)
    o = lro.pt_image(sc, W, H, 0, spp, max_depth=8, mis=True, tab=lr.brdf_tables(), mask=lro.blue_noise())
    compare(g, o, "lamp and sky: engine vs oracle")
    # the sky adds light: the image is brighter than the lamp's alone
    sc2 = dict(sc)
    del sc2["sky"]
    o2 = lro.pt_image(sc2, W, H, 0, spp, max_depth=8, mis=True, tab=lr.brdf_tables(), mask=lro.blue_noise())
    assert o.mean() > 1.05 * o2.mean()
    eng.close()
    ps.close()


def test_one_form_or_the_other_same_image_and_batched_equals_plain(ctx):
    # a lamp-lit job alone (shade and shadow with one shadow ray), then beside a job with several
    # lights (the engine switches to the other form for all jobs): the same image to the bit;
    # the several-lights job batched equals its plain form
    W = H = 40
    sc = lr.batch_scene(3000, 9)
    cam = lr.batch_camera(3, W, H)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    sky = lr.sky(128, 30.0, 120.0, sun_le=(40.0, 36.0, 30.0))
    ps2 = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    ps2.set_sky(sky)
    eng = engine(ctx, max(ps.info()[0], ps2.info()[0]))
    alone = render(eng, ps, cam, W, H, 8)
    j1 = eng.submit(lr.make_job(ps, cam, W, H, 8))
    j2 = eng.submit(lr.make_job(ps2, cam, W, H, 8))
    eng.drain()
    outs = []
    for i in (j1, j2):
        t = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        eng.image(i, t)
        eng.release(i)
        outs.append(t.cpu().numpy())
    assert np.array_equal(alone, outs[0]), "a lamp-lit image changed beside a job with several lights"
    plain = render(eng, ps2, cam, W, H, 8, plain=True)
    assert np.array_equal(plain, outs[1]), "several lights: batched and plain differ"
    st = eng.stats()
    say(f"  rays: {st.rays_ext} closest, {st.rays_shadow} shadow")
    eng.close()
    ps.close()
    ps2.close()
