"""GPU parity of the incremental updates of an instanced scene (the serving pool's ingestion
path; the reference's upload_instances, engine/gpu/trace_gpu.c:1385-1512):

- lrh_pt_scene_move: only the moved instances' records sent, scattered on the device, the top
  level rebuilt. A scene moved so equals a scene made with the final poses: the same hits on
  random rays, bit for bit, and bit-identical images from the engine. Moving nothing sends
  nothing; bad indices are refused.
- lrh_pt_scene_mesh: a dynamic mesh's new triangles, its tree rebuilt in place, the top level
  rebuilt. The scene equals one made with the new triangles, by the same checks; a wrong
  triangle count is refused.
- Either way, the updated scene's image equals the oracle's image of the final scene (brute force
  over every instance, the final poses and triangles) sample for sample, as a scene made anew
  does (test_gpu_sky_inst.py).
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


def scene_parts(seed, blob_seed=4):
    """a few meshes (a box, a sphere, a hull, a blob, the ground box) and 14 instances of them and
    of the unit sphere and box over the ground"""
    rng = np.random.default_rng(seed)
    meshes = [lr.box_mesh(1), lr.uv_sphere(12, 7, 2), lr.hull(24, 3, 3), lr.blob(300, blob_seed, 4), lr.box_mesh(5)]
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
        inst[i]["xform"] = lr.pose(np.random.default_rng([seed, i]), (s, s, s),
                                   (rng.uniform(-4, 4), s, rng.uniform(-4, 4))).reshape(12)
        inst[i]["kind"] = kinds[i % len(kinds)]
        inst[i]["material"] = int(rng.integers(0, 8))
    cam = lr.camera_make((0.0, 6.0, -12.0), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, 1.0)
    return mtris, first, mats, inst, cam


def rays_at(n, seed):
    """rays from around the scene toward points over the ground"""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-8, 8, size=(n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(1, 8, n)
    t = rng.uniform(-4, 4, size=(n, 3)).astype(np.float32)
    t[:, 1] = rng.uniform(0, 2, n)
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7], r[:, 7] = o, 1e-4, d, 1e30
    return r


def image(ctx, ps, cam, W=32, H=32, spp=4):
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=ps.info()[0], image_bytes=16 << 20)
    j = lr.make_job(ps, cam, W, H, spp, seed=3)
    i = eng.submit(j)
    eng.drain()
    t = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    eng.image(i, t)
    eng.release(i)
    eng.close()
    return t.cpu().numpy()


SKY = None


def sky():
    global SKY
    if SKY is None:
        SKY = lr.sky(64, 35.0, 60.0)
    return SKY


def oracle_image(mtris, first, mats, inst, cam, W=32, H=32, spp=4):
    """the oracle's image of the scene (the sums of samples 0 .. spp-1, seed 3), as image() takes it"""
    oi = np.zeros(len(inst), lro.INST_DTYPE)
    for k in range(len(inst)):
        oi[k]["inv"], oi[k]["kind"], oi[k]["material"] = lr.affine_inverse(inst[k]["xform"]), inst[k]["kind"], inst[k]["material"]
    sc = {"mats": mats, "light": -1, "eps": 1e-4, "margin": 1e-3, "meshes": (mtris, first), "instances": oi,
          "fwd": np.ascontiguousarray(inst["xform"]), "sky": lro.Env(sky()), "cam": cam}
    return lro.pt_image(sc, W, H, 0, spp, seed=3, max_depth=8, tab=lr.brdf_tables(), mask=lro.blue_noise())


def against_oracle(ctx, a, cam, parts, what):
    g = image(ctx, a, cam)
    o = oracle_image(*parts, cam)
    rel = np.abs(g[..., :3] - o) / np.maximum(np.abs(o), 1e-3)
    say(f"  {what} vs the oracle's final scene: median rel {np.median(rel):.2e}, >1e-3 in {np.mean(rel > 1e-3):.4f}")
    assert np.median(rel) < 2e-5
    assert np.mean(rel > 1e-3) < 0.02
    assert abs(g[..., :3].mean() - o.mean()) < 2e-3 * o.mean()


def same_scene(ctx, a, b, cam, what):
    rays = rays_at(20000, 8)
    ha, hb = a.trace(rays), b.trace(rays)
    hit = ha[:, 1].view(np.uint32) != 0xFFFFFFFF
    say(f"  {what}: {hit.sum()} of {len(rays)} rays hit")
    assert hit.mean() > 0.3
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    assert np.array_equal(a.trace(rays, any_hit=True)[:, 1].view(np.uint32) != 0xFFFFFFFF,
                          b.trace(rays, any_hit=True)[:, 1].view(np.uint32) != 0xFFFFFFFF)
    assert a.info()[0] == b.info()[0]
    ga, gb = image(ctx, a, cam), image(ctx, b, cam)
    assert ga[..., :3].mean() > 0.01
    assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))


def test_move_equals_a_scene_made_with_the_final_poses(ctx):
    mtris, first, mats, inst0, cam = scene_parts(11)
    inst1 = inst0.copy()
    idx = np.array([2, 5, 7, 11, 13], np.int32)
    rng = np.random.default_rng(99)
    for i in idx:
        inst1[i]["xform"] = lr.pose(np.random.default_rng([99, int(i)]), (0.7, 0.7, 0.7),
                                    (rng.uniform(-4, 4), 1.2, rng.uniform(-4, 4))).reshape(12)
    inst1[7]["material"] = 3 if inst0[7]["material"] != 3 else 4      # a new material override too
    say("move 5 of 14 instances")
    a = lr.PtScene.instanced(ctx, mtris, first, mats, inst0)
    a.set_sky(sky())
    st = a.move(idx, inst1[idx])
    assert st.depth >= 1
    b = lr.PtScene.instanced(ctx, mtris, first, mats, inst1)
    b.set_sky(sky())
    same_scene(ctx, a, b, cam, "moved")
    against_oracle(ctx, a, cam, (mtris, first, mats, inst1), "moved")
    # nothing moved: nothing sent, the same scene
    st = a.move(np.zeros(0, np.int32), inst1[:0])
    assert st.depth == 0 and st.ms == 0.0
    same_scene(ctx, a, b, cam, "moved, then nothing")
    with pytest.raises(lr.LrhError):
        a.move(np.array([3, 3], np.int32), inst1[[3, 3]])
    with pytest.raises(lr.LrhError):
        a.move(np.array([14], np.int32), inst1[:1])


def test_mesh_update_equals_a_scene_made_with_the_new_triangles(ctx):
    mtris0, first, mats, inst, cam = scene_parts(12)
    mtris1, first1, _, _, _ = scene_parts(12, blob_seed=9)
    assert np.array_equal(first, first1) and not np.array_equal(mtris0, mtris1)
    f, l = first[3], first[4]
    say(f"mesh 3 ({l - f} triangles) replaced")
    a = lr.PtScene.instanced(ctx, mtris0, first, mats, inst)
    a.set_sky(sky())
    st = a.mesh(3, mtris1[f:l])
    assert st.depth >= 1
    b = lr.PtScene.instanced(ctx, mtris1, first, mats, inst)
    b.set_sky(sky())
    same_scene(ctx, a, b, cam, "new mesh")
    against_oracle(ctx, a, cam, (mtris1, first, mats, inst), "new mesh")
    with pytest.raises(lr.LrhError):
        a.mesh(3, mtris1[f:l - 1])


def leaf_boxes(nodes, aux):
    """instance id -> its leaf record's box (lo.xyz, hi.xyz as float32)"""
    f = nodes.view(np.float32)
    leaf = (nodes[:, 7] & 0x80000000) != 0
    out = {}
    for r in np.nonzero(leaf)[0]:
        for k in range(int(nodes[r, 7] & 0x7FFFFFFF)):
            out[int(aux[nodes[r, 3] + k])] = np.concatenate([f[r, 0:3], f[r, 4:7]])
    return out


def cost_of(nodes):
    f = nodes.view(np.float32).astype(np.float64)
    d = f[:, 4:7] - f[:, 0:3]
    ok = (d >= 0).all(1)
    a = np.where(ok, d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0], 0.0)
    return a.sum() / a[0]


def test_refit_keeps_the_tree_while_it_stays_good(ctx):
    mtris, first, mats, inst0, cam = scene_parts(13)
    a = lr.PtScene.instanced(ctx, mtris, first, mats, inst0)
    # a small move: every instance but the ground shifted a little
    inst1 = inst0.copy()
    rng = np.random.default_rng(5)
    for i in range(1, len(inst1)):
        x = inst1[i]["xform"].reshape(3, 4).copy()
        x[:, 3] += rng.uniform(-0.05, 0.05, 3)
        inst1[i]["xform"] = x.reshape(12)
    st = a.update(inst1, builder=lr.REFIT)
    say(f"refit: refitted {st.refitted}, {st.ms:.3f} ms")
    assert st.refitted == 1 and st.depth >= 1
    b = lr.PtScene.instanced(ctx, mtris, first, mats, inst1)
    na, xa, ca = a.top()
    nb, xb, cb = b.top()
    # every inner record's box is its children's union, every leaf's its instance's world box
    f = na.view(np.float32)
    inner = np.nonzero((na[:, 7] & 0x80000000) == 0)[0]
    for r in inner:
        lc, rc = na[r, 3], na[r, 7]
        assert np.array_equal(f[r, 0:3], np.minimum(f[lc, 0:3], f[rc, 0:3]))
        assert np.array_equal(f[r, 4:7], np.maximum(f[lc, 4:7], f[rc, 4:7]))
    la, lb = leaf_boxes(na, xa), leaf_boxes(nb, xb)
    assert sorted(la) == list(range(len(inst1)))
    for i in la:
        assert np.array_equal(la[i], lb[i]), i
    assert abs(ca - cost_of(na)) <= 1e-9 * ca and ca >= 1.0
    # the same hits as the tree built anew (a tie, two shapes at one distance, may differ)
    rays = rays_at(20000, 9)
    ha, hb = a.trace(rays), b.trace(rays)
    ia, ib = ha[:, 1].view(np.uint32), hb[:, 1].view(np.uint32)
    assert np.array_equal(ia == 0xFFFFFFFF, ib == 0xFFFFFFFF)
    same = (ia == ib) & (ia != 0xFFFFFFFF)
    assert np.array_equal(ha[same].view(np.uint32), hb[same].view(np.uint32))
    diff = (ia != ib)
    assert diff.mean() < 0.001 and np.all(np.abs(ha[diff, 0] - hb[diff, 0]) <= 1e-5 * np.abs(hb[diff, 0]))
    # a large move (the instances' places shuffled): the cost grows past 1.1, the tree is built
    # again with the scene's builder, as a scene made with these poses has it
    inst2 = inst1.copy()
    perm = np.r_[0, 1 + rng.permutation(len(inst2) - 1)]
    x = inst2["xform"].reshape(-1, 3, 4).copy()
    x[:, :, 3] = x[perm, :, 3]
    inst2["xform"] = x.reshape(-1, 12)
    st = a.update(inst2, builder=lr.REFIT)
    say(f"large move: refitted {st.refitted}")
    assert st.refitted == 0
    c = lr.PtScene.instanced(ctx, mtris, first, mats, inst2)
    n2, x2, c2 = a.top()
    nc_, xc, cc = c.top()
    assert np.array_equal(n2, nc_) and np.array_equal(x2, xc) and c2 == cc
