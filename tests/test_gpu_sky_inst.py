"""GPU parity of the sky and of two-level (instanced) scenes.

- A scene lit by the sky alone (escaping rays see it, next event estimation samples it with MIS)
  against the oracle sample for sample: the device's atan2 / acos / sin / cos may round
  differently from glibc's in the last bit, so a sample can land in the neighbouring texel;
  the median pixel agrees to ~1e-5 and the image mean to 1e-3.
- An instanced scene (mesh instances, unit spheres and boxes, per-instance materials) against the
  oracle's brute force over every instance, sample for sample; after new poses (the top level
  rebuilt on the device) again.
- Hits of an instanced scene against the oracle's brute force on the same rays: the same
  instance and triangle, t within 1e-5 relative (ties aside).
- The engine's equal-error guarantee holds for these scenes too: plain and batched images are
  bit-identical, alone and in a mix with lamp-lit scenes.
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


SKY = None


def sky():
    global SKY
    if SKY is None:
        SKY = lr.sky(256, 35.0, 60.0, sun_le=(5e4, 4.6e4, 4e4))
    return SKY


def render(eng, ps, cam, W, H, spp, seed=0, max_depth=8, plain=False):
    j = lr.make_job(ps, cam, W, H, spp, seed=seed, max_depth=max_depth)
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
    assert abs(g[..., :3].mean() - o.mean()) < 2e-3 * o.mean()


def sky_scene():
    sc = lr.batch_scene(1500, 21)
    sc["prims"] = np.zeros((0, 24), np.float32)      # no lamp: the sky lights it
    sc["light"] = -1
    return sc


def test_sky_scene_matches_oracle(ctx):
    W = H = 40
    sc = sky_scene()
    sc["cam"] = lr.batch_camera(77, W, H)
    ps = lr.PtScene(ctx, None, sc["tris"], sc["mats"], -1, sc["eps"], sc["margin"])
    ps.set_sky(sky())
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=ps.info()[0], image_bytes=16 << 20)
    say("sky scene vs oracle 40x40x8")
    g = render(eng, ps, sc["cam"], W, H, 8, seed=5)
    eng.close()
    osc = dict(sc, sky=lro.Env(sky()))
    o = lro.pt_image(osc, W, H, 0, 8, seed=5, tab=lr.brdf_tables(), mask=lro.blue_noise())
    assert o.mean() > 0.05
    compare(g, o, "sky")


def small_instanced(seed, frame=0):
    """a few meshes (a sphere, a box, a hull, a blob), instances of each and of the unit sphere and
    box, each with a material or the mesh's own, over a ground instance"""
    rng = np.random.default_rng(seed)
    meshes = [lr.box_mesh(1), lr.uv_sphere(12, 7, 2), lr.hull(24, 3, 3), lr.blob(300, 4, 4)]
    ground = lr.box_mesh(5)
    meshes.append(ground)
    first = np.cumsum([0] + [len(m) for m in meshes]).astype(np.int32)
    mtris = np.concatenate(meshes)
    mats = (lr.Material * 8)()
    bs = lr.batch_scene(100, seed)["mats"]
    for k in range(8):
        mats[k] = bs[k]
    n = 14
    inst = np.zeros(n, lr.INSTANCE_DTYPE)
    inst[0]["xform"] = np.array([6, 0, 0, 0, 0, 0.1, 0, -0.1, 0, 0, 6, 0], np.float32)
    inst[0]["kind"], inst[0]["material"] = 4, -1
    kinds = [0, 1, 2, 3, lr.INST_SPHERE, lr.INST_BOX]
    for i in range(1, n):
        s = rng.uniform(0.4, 1.0)
        inst[i]["xform"] = lr.pose(np.random.default_rng([seed, i, frame]), (s, s * rng.uniform(0.7, 1.3), s),
                                   (rng.uniform(-4, 4), s + 0.3 * frame, rng.uniform(-4, 4))).reshape(12)
        inst[i]["kind"] = kinds[i % len(kinds)]
        inst[i]["material"] = -1 if i % 4 == 0 and kinds[i % len(kinds)] >= 0 else int(rng.integers(0, 8))
    cam = lr.camera_make((0.0, 6.0, -12.0), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, 1.0)
    return mtris, first, mats, inst, cam


def oracle_inst(mtris, first, mats, inst):
    oi = np.zeros(len(inst), lro.INST_DTYPE)
    for k in range(len(inst)):
        oi[k]["inv"] = lr.affine_inverse(inst[k]["xform"])
        oi[k]["kind"], oi[k]["material"] = inst[k]["kind"], inst[k]["material"]
    return {"mats": mats, "eps": 1e-4, "light": -1, "meshes": (mtris, first), "instances": oi}


def test_instanced_hits_match_oracle(ctx):
    mtris, first, mats, inst, cam = small_instanced(3)
    ps = lr.PtScene.instanced(ctx, mtris, first, mats, inst)
    osc = oracle_inst(mtris, first, mats, inst)
    # rays from above aimed near the instances' centres, through the scene's own walk
    rng = np.random.default_rng(1)
    n = 3000
    o = rng.uniform(-8, 8, size=(n, 3)).astype(np.float32)
    o[:, 1] = np.abs(o[:, 1]) + 1
    at = np.array([np.asarray(inst[k]["xform"]).reshape(3, 4)[:, 3] for k in range(1, len(inst))])
    tgt = at[rng.integers(0, len(at), n)] + rng.normal(scale=0.3, size=(n, 3))
    d = (tgt - o)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 1e-4, d, 1e30
    want = lro.scene_trace(osc, rays)
    got = ps.trace(rays)
    wi, gi = want[:, 1].view(np.uint32), got[:, 1].view(np.uint32)
    winst, ginst = want[:, 2].view(np.uint32), got[:, 2].view(np.uint32)
    hit = wi != 0xffffffff
    assert hit.mean() > 0.5
    assert np.array_equal(hit, gi != 0xffffffff)
    same_inst = winst[hit] == ginst[hit]
    # the device's triangle ids are global (the mesh's first + local)
    kinds = np.asarray(inst["kind"])
    g_local = gi[hit].astype(np.int64)
    tri = kinds[ginst[hit]] >= 0
    g_local[tri] -= first[kinds[ginst[hit]][tri]]
    same = same_inst & (g_local.astype(np.uint32) == wi[hit])
    rel = np.abs(got[hit, 0] - want[hit, 0]) / want[hit, 0]
    say(f"  instanced hits: {hit.sum()} hits, same (instance, shape) {same.mean():.4f}, 99.9% rel t {np.percentile(rel, 99.9):.2e}")
    assert same.mean() > 0.998
    assert np.percentile(rel[same], 99.9) < 1e-5


def test_instanced_scene_matches_oracle(ctx):
    W = H = 32
    mtris, first, mats, inst, cam = small_instanced(5)
    ps = lr.PtScene.instanced(ctx, mtris, first, mats, inst)
    ps.set_sky(sky())
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=ps.info()[0], image_bytes=16 << 20)
    say("instanced scene vs oracle 32x32x4")
    g = render(eng, ps, cam, W, H, 4, seed=2)
    osc = dict(oracle_inst(mtris, first, mats, inst), sky=lro.Env(sky()), cam=cam)
    o = lro.pt_image(osc, W, H, 0, 4, seed=2, tab=lr.brdf_tables(), mask=lro.blue_noise())
    compare(g, o, "instanced")
    # new poses: the top level rebuilt on the device
    mtris2, first2, mats2, inst2, _ = small_instanced(5, frame=1)
    st = ps.update(inst2)
    assert st.depth >= 1
    g2 = render(eng, ps, cam, W, H, 4, seed=2)
    osc2 = dict(oracle_inst(mtris, first, mats, inst2), sky=lro.Env(sky()), cam=cam)
    o2 = lro.pt_image(osc2, W, H, 0, 4, seed=2, tab=lr.brdf_tables(), mask=lro.blue_noise())
    compare(g2, o2, "instanced, moved")
    assert not np.allclose(g2, g)
    # plain and batched, alone and beside a lamp-lit job: bit-identical
    cb = lr.cornell(W, H)
    pc = lr.PtScene(ctx, cb["prims"], cb["tris"], cb["mats"], 0, cb["eps"], cb["margin"])
    jobs = [lr.make_job(ps, cam, W, H, 4, seed=2), lr.make_job(pc, cb["cam"], W, H, 4, seed=9)]
    ids = [eng.submit(j) for j in jobs]
    eng.drain()
    mix = []
    for i, j in zip(ids, jobs):
        t = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        eng.image(i, t)
        eng.release(i)
        mix.append(t.cpu().numpy())
    plain = render(eng, ps, cam, W, H, 4, seed=2, plain=True)
    eng.close()
    assert np.array_equal(mix[0].view(np.uint32), g2.view(np.uint32))
    assert np.array_equal(plain.view(np.uint32), g2.view(np.uint32))
