"""GPU tests of the engine (many jobs over one pool of path slots) and of the serving loop.

A job's samples depend only on (its seed, pixel, sample index), and each sample joins its pixel as
a fixed-point integer, so a job's image must be the same to the bit whatever else runs beside it,
however the jobs are launched (the plain per-job form, the batched graph, any launch shape) and
however few slots the engine has. These tests check exactly that (bit-identical images), that the
engine's image equals the oracle's sample for sample, and that serving completes every request
with the same images under a memory budget that admits jobs one at a time.
"""
import sys
import threading

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


@pytest.fixture(scope="module")
def scenes(ctx):
    """the Cornell box and three small scenes of the batched workload"""
    out = []
    cb = lr.cornell(64, 64)
    out.append((lr.PtScene(ctx, cb["prims"], cb["tris"], cb["mats"], 0, cb["eps"], cb["margin"]), cb))
    for n, seed in ((1000, 11), (6000, 12), (30000, 13)):
        sc = lr.batch_scene(n, seed)
        out.append((lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], 0, sc["eps"], sc["margin"]), sc))
    yield out
    for ps, _ in out:
        ps.close()


def job_set(scenes):
    """a mix of scenes, resolutions, sample counts, seeds and depths"""
    jobs = []
    specs = [(0, 64, 64, 8, 0, 8), (1, 96, 40, 4, 3, 8), (2, 50, 70, 16, 5, 6), (3, 128, 128, 2, 9, 8),
             (0, 33, 17, 5, 2, 3), (2, 200, 120, 1, 7, 8), (1, 64, 64, 3, 1, 0)]
    for k, (si, w, h, spp, seed, depth) in enumerate(specs):
        ps, sc = scenes[si]
        cam = lr.cornell(w, h)["cam"] if si == 0 else lr.batch_camera(100 + k, w, h)
        jobs.append(lr.make_job(ps, cam, w, h, spp, s0=k, seed=seed, max_depth=depth, mis=(k % 3 != 2)))
    return jobs


def images(eng, ids, jobs):
    out = []
    for i, j in zip(ids, jobs):
        t = torch.zeros((j.h, j.w, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        eng.image(i, t)
        out.append(t.cpu().numpy())
        eng.release(i)
    return out


def run_batched(eng, jobs):
    ids = [eng.submit(j) for j in jobs]
    assert all(i is not None for i in ids)
    eng.drain()
    return images(eng, ids, jobs)


def run_plain(eng, jobs):
    return images(eng, [eng.plain(j) for j in jobs], jobs)


def same(a, b):
    return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))


def test_engine_matches_oracle_sample_for_sample(ctx, scenes):
    ps, cb = scenes[0]
    say("engine: Cornell 48x48x16 vs oracle")
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=64, image_bytes=64 << 20)
    job = lr.make_job(ps, lr.cornell(48, 48)["cam"], 48, 48, 16)
    g = run_batched(eng, [job])[0]
    eng.close()
    assert np.all(g[..., 3] == 16)
    o = lro.pt_image(lr.cornell(48, 48), 48, 48, 0, 16, tab=lr.brdf_tables(), mask=lro.blue_noise())
    rel = np.abs(g[..., :3] - o) / np.maximum(np.abs(o), 1e-3)
    say(f"  median rel {np.median(rel):.2e}, max {rel.max():.2e}")
    assert np.median(rel) < 1e-5 and np.mean(rel > 1e-3) < 0.01


def test_engine_lamp_lit_ggx_scene_matches_oracle(ctx):
    # a scene of the batched workload as the benchmark renders it: a 30,000-triangle soup of 14
    # GGX materials (metals, rough dielectrics, two clear glasses), a diffuse floor, the quad
    # lamp; two cameras, 64 x 64, 16 spp, against the oracle sample for sample
    sc = lr.batch_scene(30000, 31)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], 0, sc["eps"], sc["margin"])
    W = H = 64
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=max(64, ps.info()[0]), image_bytes=64 << 20)
    for cam_seed in (301, 302):
        cam = lr.batch_camera(cam_seed, W, H)
        say(f"engine: lamp-lit GGX soup 30000 tris, camera {cam_seed}, {W}x{H}x16 vs oracle")
        g = run_batched(eng, [lr.make_job(ps, cam, W, H, 16, seed=7)])[0]
        assert np.all(g[..., 3] == 16)
        o = lro.pt_image(dict(sc, cam=cam), W, H, 0, 16, seed=7, tab=lr.brdf_tables(), mask=lro.blue_noise())
        rel = np.abs(g[..., :3] - o) / np.maximum(np.abs(o), 1e-3)
        say(f"  median rel {np.median(rel):.2e}, >1e-3 in {np.mean(rel > 1e-3):.4f}, means {g[..., :3].mean():.5f} "
            f"{o.mean():.5f}")
        assert o.mean() > 0
        assert np.median(rel) < 2e-5
        assert np.mean(rel > 1e-3) < 0.02
        assert abs(g[..., :3].mean() - o.mean()) < 2e-3 * o.mean()
    eng.close()
    ps.close()


def test_server_jobs_match_the_oracle(ctx):
    # the serving path end to end: a resident scene, jobs submitted from client threads, each
    # job's image (read back through the server) against the oracle sample for sample
    sc = lr.batch_scene(6000, 33)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], 0, sc["eps"], sc["margin"])
    srv = lr.Server(ctx, slots_log2=16, max_tree_depth=max(64, ps.info()[0]), budget=1 << 30)
    srv.add_scene(ps)
    W, H = 48, 40
    specs = [(401, 8, 3), (402, 4, 9), (403, 12, 1)]
    got = [None] * len(specs)

    def client(k):
        cam_seed, spp, seed = specs[k]
        i = srv.submit(lr.make_job(ps, lr.batch_camera(cam_seed, W, H), W, H, spp, seed=seed))
        info = srv.wait(i, 120000.0)
        assert info is not None and info.state == lr.JOB_DONE
        got[k] = srv.image(i, W, H)
        srv.release(i)

    th = [threading.Thread(target=client, args=(k,)) for k in range(len(specs))]
    for t in th:
        t.start()
    for t in th:
        t.join(180)
    srv.close()
    tab, mask = lr.brdf_tables(), lro.blue_noise()
    for (cam_seed, spp, seed), g in zip(specs, got):
        assert g is not None and np.all(g[..., 3] == spp)
        o = lro.pt_image(dict(sc, cam=lr.batch_camera(cam_seed, W, H)), W, H, 0, spp, seed=seed, tab=tab, mask=mask)
        rel = np.abs(g[..., :3] - o) / np.maximum(np.abs(o), 1e-3)
        say(f"server job (camera {cam_seed}, {spp} spp): median rel {np.median(rel):.2e}, >1e-3 in {np.mean(rel > 1e-3):.4f}")
        assert o.mean() > 0 and np.median(rel) < 2e-5 and np.mean(rel > 1e-3) < 0.02
        assert abs(g[..., :3].mean() - o.mean()) < 2e-3 * o.mean()
    ps.close()


def test_batched_equals_plain_bit_for_bit(ctx, scenes):
    jobs = job_set(scenes)
    say(f"engine: {len(jobs)} jobs batched vs plain")
    eng = lr.Engine(ctx, slots_log2=17, max_tree_depth=64, image_bytes=256 << 20)
    a = run_batched(eng, jobs)
    s1 = eng.stats()
    b = run_plain(eng, jobs)
    s2 = eng.stats()
    eng.close()
    assert same(a, b)
    for j, im in zip(jobs, a):
        assert np.all(im[..., 3] == j.spp)
        assert np.isfinite(im).all()
        assert j.max_depth == 0 or im[..., :3].sum() > 0   # depth 0: only a lamp seen directly
    # the same paths: the same rays, counted on the device
    assert s2.rays_ext - s1.rays_ext == s1.rays_ext and s2.rays_shadow - s1.rays_shadow == s1.rays_shadow
    say(f"  rays {s1.rays_ext} + {s1.rays_shadow} shadow")


def test_alone_and_in_the_mix(ctx, scenes):
    jobs = job_set(scenes)
    say("engine: each job alone vs all together")
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=64, image_bytes=256 << 20)
    mix = run_batched(eng, jobs)
    alone = [run_batched(eng, [j])[0] for j in jobs]
    # and in another order, the pool small enough that paths of many jobs share every iteration
    rev = run_batched(eng, jobs[::-1])[::-1]
    eng.close()
    assert same(mix, alone)
    assert same(mix, rev)


def test_few_slots_and_every_launch_shape(ctx, scenes):
    jobs = job_set(scenes)[:4]
    say("engine: launch shapes and a pool of 1024 slots")
    big = lr.Engine(ctx, slots_log2=18, max_tree_depth=64, image_bytes=128 << 20)
    ref = run_batched(big, jobs)
    big.close()
    small = lr.Engine(ctx, slots_log2=10, max_tree_depth=64, image_bytes=128 << 20)
    assert same(ref, run_batched(small, jobs))
    small.close()
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=64, image_bytes=128 << 20)
    # refill: the walks handing new rays to finished lanes (1: at once, 64: a whole wave's at a
    # time) or not (0); with few persistent waves (one work-group of 64 a compute unit) each wave
    # takes many runs of the queue, so that lanes are refilled in the middle of others' walks
    shapes = [dict(block_trace=64, stack=16, block_shade=256, blocks_per_cu=1, tile=1, iters_per_graph=1, refill=4,
                   refill_shadow=4),
              dict(block_trace=256, stack=32, block_shade=64, blocks_per_cu=16, tile=8, iters_per_graph=64, refill=1,
                   refill_shadow=0),
              dict(block_trace=128, stack=24, block_shade=128, blocks_per_cu=2, tile=4, iters_per_graph=3, refill=64,
                   refill_shadow=64),
              dict(block_trace=256, stack=16, block_shade=256, blocks_per_cu=8, tile=1, iters_per_graph=8, refill=16,
                   refill_shadow=16),
              dict(block_trace=64, stack=24, block_shade=64, blocks_per_cu=4, tile=2, iters_per_graph=2, refill=0,
                   refill_shadow=1)]
    for t in shapes:
        eng.tune = t
        assert same(ref, run_batched(eng, jobs)), t
        assert same(ref, run_plain(eng, jobs)), t
    eng.close()


def test_autotune_keeps_images(ctx, scenes):
    jobs = job_set(scenes)
    say("engine: autotune")
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=64, image_bytes=256 << 20)
    ref = run_batched(eng, jobs)
    sample = [lr.make_job(scenes[3][0], lr.batch_camera(7, 128, 128), 128, 128, 1, seed=4)]
    best, mrays = eng.autotune(sample)
    say(f"  best {best.as_dict()} at {mrays:.1f} Mrays/s")
    assert mrays > 0 and eng.tune.as_dict() == best.as_dict()
    assert same(ref, run_batched(eng, jobs))
    eng.close()


def test_counting_walks(ctx, scenes):
    jobs = job_set(scenes)[:3]
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=64, image_bytes=64 << 20)
    # few persistent waves (one of 64 lanes a compute unit), so that refilled lanes take many runs
    eng.tune = dict(block_trace=64, blocks_per_cu=1, refill=0, refill_shadow=0)
    ref = run_batched(eng, jobs)
    eng.counting(True)
    got = run_batched(eng, jobs)
    st = eng.stats()
    # the walk with lanes refilled visits the same nodes and shapes (only its waves' steps differ)
    eng.tune = dict(refill=8, refill_shadow=8)
    got_rf = run_batched(eng, jobs)
    st_rf = eng.stats()
    eng.close()
    assert same(ref, got) and same(ref, got_rf)
    d = lambda f: getattr(st_rf, f) - getattr(st, f)
    assert d("node_fetches") == st.node_fetches and d("shape_tests") == st.shape_tests
    assert d("shadow_node_fetches") == st.shadow_node_fetches and d("shadow_shape_tests") == st.shadow_shape_tests
    say(f"  refilled lanes: extend node steps {st.node_fetches / (64 * d('wave_node_steps')):.3f} (against "
        f"{st.node_fetches / (64 * st.wave_node_steps):.3f}), shape steps "
        f"{st.shape_tests / (64 * d('wave_shape_steps')):.3f}")
    assert st.node_fetches > st.rays_ext // 2 and st.shape_tests > 0
    # every wave step has between 1 and 64 lanes at work in it, in both walks
    for n, w in ((st.node_fetches, st.wave_node_steps), (st.shape_tests, st.wave_shape_steps),
                 (st.shadow_node_fetches, st.shadow_wave_node_steps), (st.shadow_shape_tests, st.shadow_wave_shape_steps)):
        assert 0 < w <= n <= 64 * w
    say(f"  lanes at work: extend node steps {st.node_fetches / (64 * st.wave_node_steps):.3f}, "
        f"shape steps {st.shape_tests / (64 * st.wave_shape_steps):.3f}; shadow node steps "
        f"{st.shadow_node_fetches / (64 * st.shadow_wave_node_steps):.3f}, "
        f"shape steps {st.shadow_shape_tests / (64 * st.shadow_wave_shape_steps):.3f}")
    # a shadow ray fetches at least its root
    assert st.shadow_node_fetches >= st.rays_shadow
    # uniform node steps (every lane at one node) are node steps, and the roots' are among them
    assert 0 < st.uniform_node_steps <= st.wave_node_steps and 0 < st.shadow_uniform_node_steps <= st.shadow_wave_node_steps


def test_serving_same_results_under_a_tight_budget(ctx, scenes):
    jobs = job_set(scenes)
    say("serving: threads submit, a budget that fits one image at a time")
    eng = lr.Engine(ctx, slots_log2=16, max_tree_depth=64, image_bytes=256 << 20)
    ref = run_batched(eng, jobs)
    eng.close()
    biggest = max(j.w * j.h * 24 for j in jobs)
    srv = lr.Server(ctx, slots_log2=16, max_tree_depth=64, budget=1 << 30)
    for ps, _ in scenes:
        srv.add_scene(ps)
    st0 = srv.stats()
    srv.close()
    # the budget: the scenes plus room for the largest image (aligned) and a little less than two
    budget = st0.in_use + biggest + 4096
    srv = lr.Server(ctx, slots_log2=16, max_tree_depth=64, budget=budget)
    for ps, _ in scenes:
        srv.add_scene(ps)
    ids = [None] * len(jobs)
    got = [None] * len(jobs)

    def client(k):
        ids[k] = srv.submit(jobs[k])
        info = srv.wait(ids[k], 120000.0)
        assert info is not None and info.state == lr.JOB_DONE
        assert info.t_done_ms >= info.t_submit_ms
        got[k] = srv.image(ids[k], jobs[k].w, jobs[k].h)
        srv.release(ids[k])

    th = [threading.Thread(target=client, args=(k,)) for k in range(len(jobs))]
    for t in th:
        t.start()
    for t in th:
        t.join(180)
    st = srv.stats()
    srv.close()
    say(f"  done {st.done}, p50 {st.p50_ms:.2f} ms, p99 {st.p99_ms:.2f} ms, rays {st.rays}")
    assert st.done == len(jobs) and st.submitted == len(jobs)
    assert st.p99_ms >= st.p50_ms > 0
    assert same(ref, got)
