"""GPU parity of the incremental updates of an instanced scene (the serving pool's ingestion
path; the reference's upload_instances, engine/gpu/trace_gpu.c:1385-1512):

- lrh_pt_scene_move: only the moved instances' records sent, scattered on the device, the top
  level rebuilt. A scene moved so equals a scene made with the final poses: the same hits on
  random rays, bit for bit, and bit-identical images from the engine. Moving nothing sends
  nothing; bad indices are refused.
- lrh_pt_scene_mesh: a dynamic mesh's new triangles, its tree rebuilt in place, the top level
  rebuilt. The scene equals one made with the new triangles, by the same checks; a wrong
  triangle count is refused.
"""
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
    a.set_sky(lr.sky(64, 35.0, 60.0))
    st = a.move(idx, inst1[idx])
    assert st.depth >= 1
    b = lr.PtScene.instanced(ctx, mtris, first, mats, inst1)
    b.set_sky(lr.sky(64, 35.0, 60.0))
    same_scene(ctx, a, b, cam, "moved")
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
    a.set_sky(lr.sky(64, 35.0, 60.0))
    st = a.mesh(3, mtris1[f:l])
    assert st.depth >= 1
    b = lr.PtScene.instanced(ctx, mtris1, first, mats, inst)
    b.set_sky(lr.sky(64, 35.0, 60.0))
    same_scene(ctx, a, b, cam, "new mesh")
    with pytest.raises(lr.LrhError):
        a.mesh(3, mtris1[f:l - 1])
