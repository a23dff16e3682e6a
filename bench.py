#!/usr/bin/env python3
"""bench.py - little_renderer's path-tracing hot path on one MI355X.

Default workload (--workload batched): the metric's own form. One step = a batch of 64 render
jobs, each its own resident scene (seeded triangle soup of 1k-500k triangles, 14 GGX materials,
floor, quad lamp), camera, resolution (256^2 .. 1920x1080) and spp (1 .. 64), traced together
by the engine (persistent wavefront kernels replayed as one hipGraph, autotuned launch shape).
value = rays traced (closest-hit + shadow, counted on the device) per second over the timed
batches. Beside it in the same JSON line: the same jobs in the plain per-job form (the
reference's shape) with the images compared bit for bit, p50/p99 job latency of both, each
kernel's time from HIP events on the engine's stream, the roofline of the dominant kernel (the
extension walk: algorithmic bytes from its own counters over its time), serving under Poisson
arrivals at several rates, and the reference's CPU path tracer (compiled from its sources into
oracle/_ref) on a sample of the same jobs on this host's cores.

--workload bvh_stress: the 1M-triangle BVH-stress step (LBVH + PLOC builds, 1920x1080 primaries,
a diffuse bounce and a shadow ray per hit), with its own roofline and CPU baseline.

--workload caustics: one job with light-traced caustics (a glass and a mirror ball in the Cornell
room, 512x512, 16 spp, one light path a pixel and pass); the light kernel's time, roofline and
HBM traffic, and the reference's CPU path (trace_image with skip_caustics + light_trace_caustics).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from little_renderer_amd import roofline as rf

HBM_PEAK_GBS = rf.HBM_PEAK_GBS   # MI355X HBM3E, published


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (batched: 2 batches; bvh_stress: 10)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (batched: 1; bvh_stress: 3)")
    ap.add_argument("--workload", choices=["batched", "bvh_stress", "instanced", "caustics"], default="batched")
    ap.add_argument("--jobs", type=int, default=64, help="batched: jobs (and scenes) in a batch")
    ap.add_argument("--serve-requests", type=int, default=200, help="serving: requests per arrival rate (0: skip)")
    ap.add_argument("--serve-rates", default="10,25,50,100,200,400",
                    help="serving: Poisson arrival rates, jobs/s (the workload's 50-400, and two below saturation)")
    ap.add_argument("--tune", default=None, metavar="BT,STK,BS,BPC,TILE,IPG[,REFILL[,REFILL_SHADOW]]",
                    help="batched: this launch shape instead of the autotuner's")
    ap.add_argument("--skip-plain", action="store_true", help="batched: leave out the plain per-job comparison")
    ap.add_argument("--tlas", choices=["lbvh", "ploc", "refit"], default="ploc",
                    help="instanced: the per-frame top-level builder (refit: the same tree while its cost stays within 1.1)")
    ap.add_argument("--frames-in-flight", type=int, default=3, help="instanced: frames rendering at once")
    ap.add_argument("--physics", action="store_true",
                    help="instanced: the dynamic instances moved by the rigid-body step (1/60 s a frame, 4 substeps) "
                         "instead of the workload's scripted poses")
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--build-sizes", default="4000000,16000000",
                    help="bvh_stress: also time the builds alone at these sizes (the build kernels' roofline at scale)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-traffic", action="store_true",
                    help="batched: skip the extension kernel's HBM traffic (rocprofv3 --pmc child runs)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline: repeat for about this long")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 3 times (calibrates the PMC byte counters) and exit")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (the work would shard by "
              f"resident scene or image tile across GPUs; only the one-GPU path is built)", file=sys.stderr)
        sys.exit(2)
    import torch
    import little_renderer_amd as lr

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    if a.copy_probe:
        src = torch.ones(a.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.copy_probe << 20, "launches": 3}), flush=True)
        return
    if a.workload == "bvh_stress":
        out = bench_bvh_stress(a, torch, lr)
    elif a.workload == "instanced":
        out = bench_instanced(a, torch, lr)
    elif a.workload == "caustics":
        out = bench_caustics(a, torch, lr)
    else:
        out = bench_batched(a, torch, lr)
    print(json.dumps(out), flush=True)


def bench_bvh_stress(a, torch, lr):
    """the BVH-stress workload: builds and one plain launch sequence of traces per step"""
    a.steps = a.steps or 10
    a.warmup = 3 if a.warmup is None else a.warmup
    ctx = lr.Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream)
    n, W, H = a.tris, a.width, a.height
    npix = W * H

    # ---- inputs, resident in HBM before timing ----
    tris_np = lr.gen_soup(n, 2)
    cam = lr.camera_make((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 40.0, W / H)
    light = np.array([0.3, 1.6, -0.4], np.float32)
    eps = 1e-4
    dev = "cuda"
    tris = torch.from_numpy(tris_np).to(dev)
    f32, i32 = torch.float32, torch.int32
    boxes = torch.empty((n, 8), dtype=f32, device=dev)
    nodes_l = torch.empty((2 * n - 1, 8), dtype=i32, device=dev)
    aux_l = torch.empty(n, dtype=i32, device=dev)
    nodes_p = torch.empty((2 * n - 1, 8), dtype=i32, device=dev)
    aux_p = torch.empty(n, dtype=i32, device=dev)
    rays0 = torch.empty((npix, 8), dtype=f32, device=dev)
    hits0 = torch.empty((npix, 4), dtype=f32, device=dev)
    bounce = torch.empty((npix, 8), dtype=f32, device=dev)
    shadow = torch.empty((npix, 8), dtype=f32, device=dev)
    pix = torch.empty(npix, dtype=i32, device=dev)
    count = torch.zeros(1, dtype=i32, device=dev)
    hitsb = torch.empty((npix, 4), dtype=f32, device=dev)
    hitss = torch.empty((npix, 4), dtype=f32, device=dev)
    torch.cuda.synchronize()
    ctx.tri_boxes(tris, boxes)
    st_l = ctx.build(boxes, n, nodes_l, aux_l, lr.LBVH)
    st_p = ctx.build(boxes, n, nodes_p, aux_p, lr.PLOC)
    depth = max(st_l.depth, st_p.depth)

    names = ["build_lbvh", "build_ploc", "rays_primary", "trace_primary", "rays_secondary", "trace_bounce",
             "trace_shadow"]
    evs = {k: [] for k in names}

    def mark(k, fn):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        fn()
        e.record(stream)
        evs[k].append((s, e))

    def step(record):
        m = mark if record else (lambda k, fn: fn())
        m("build_lbvh", lambda: ctx.build(boxes, n, nodes_l, aux_l, lr.LBVH, stats=False))
        m("build_ploc", lambda: ctx.build(boxes, n, nodes_p, aux_p, lr.PLOC, stats=False))
        m("rays_primary", lambda: ctx.rays_primary(cam, W, H, 0.0, 1e30, rays0))
        m("trace_primary", lambda: ctx.trace(nodes_p, 0, aux_p, tris, rays0, npix, hits0, depth))
        m("rays_secondary", lambda: ctx.rays_secondary(rays0, hits0, npix, tris, light, eps, 7, bounce, shadow, pix,
                                                       count))
        m("trace_bounce", lambda: ctx.trace_queue(nodes_p, 0, aux_p, tris, bounce, npix, count, hitsb, depth))
        m("trace_shadow", lambda: ctx.trace_queue(nodes_p, 0, aux_p, tris, shadow, npix, count, hitss, depth,
                                                  any_hit=True))

    for _ in range(a.warmup):
        step(False)
    ctx.sync()
    nsec = int(count.cpu().item())
    rays_per_step = npix + 2 * nsec

    # ---- the timed region: exactly K steps between two full synchronizations ----
    torch.cuda.synchronize()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step(True)
    ctx.sync()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt * 1000.0 / a.steps
    value = rays_per_step * a.steps / dt / 1e6
    kms = {k: float(np.mean([s.elapsed_time(e) for s, e in v])) for k, v in evs.items()}

    # ---- the roofline of the dominant kernel (the longest per launch): its algorithmic bytes
    # from the walk's own counters (node records fetched x 32 B, triangles tested x 52 B, the
    # ray 32 B and the hit 16 B), over its average launch time from the events above ----
    trace_kernels = {"trace_primary": (rays0, None, False), "trace_bounce": (bounce, count, False),
                     "trace_shadow": (shadow, count, True)}
    dom = max(trace_kernels, key=lambda k: kms[k])
    r, c, anyh = trace_kernels[dom]
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    scratch_hits = torch.empty((npix, 4), dtype=f32, device=dev)
    ctx.trace_count(nodes_p, 0, aux_p, tris, r, npix, c, scratch_hits, depth, counts, any_hit=anyh)
    ctx.sync()
    cn, ct = [int(x) for x in counts.cpu().tolist()]
    nr = npix if c is None else nsec
    alg_bytes = rf.walk_bytes(nr, cn, ct, rf.NODE_BYTES_BINARY)
    achieved, _ = rf.roofline(alg_bytes, kms[dom])
    lb_gbs, lb_frac = rf.roofline(rf.lbvh_bytes(n), kms["build_lbvh"])
    pl_gbs, pl_frac = rf.roofline(rf.ploc_bytes(n, st_p.clusters), kms["build_ploc"])

    out = {
        "metric": "batched Mrays/s + p50/p99 job latency vs plain per-job launch, equal RMSE",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded triangle soup, generated on the host)",
        "config": {"workload": "bvh_stress: 1M-triangle soup, LBVH + PLOC build, 1920x1080 primaries + diffuse "
                               "bounce + shadow rays (one plain launch sequence per step)",
                   "tris": n, "width": W, "height": H, "rays_per_step": rays_per_step, "secondary": nsec,
                   "tree_depth": depth},
        "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
        "build_mprims_s": {"lbvh": round(n / (kms["build_lbvh"] * 1e3), 2),
                           "ploc": round(n / (kms["build_ploc"] * 1e3), 2)},
        "build_roofline": {"kernel": "build_lbvh (bounds, morton, 4 sort passes, lbvh, lbvh_top)", "bound": "hbm",
                           "alg_bytes": rf.lbvh_bytes(n), "achieved": round(lb_gbs, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(lb_frac, 4),
                           "ploc": {"alg_bytes": rf.ploc_bytes(n, st_p.clusters), "live": st_p.clusters,
                                    "achieved": round(pl_gbs, 2), "frac": round(pl_frac, 4)}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes": alg_bytes, "node_fetches": cn, "tri_tests": ct, "rays": nr},
    }

    # ---- the builds alone at larger sizes: where launch and dependency latency stop dominating ----
    scaling = []
    for m in [int(x) for x in a.build_sizes.split(",") if x]:
        t_m = torch.from_numpy(lr.gen_soup(m, 3)).to(dev)
        b_m = torch.empty((m, 8), dtype=f32, device=dev)
        nd_m = torch.empty((2 * m - 1, 8), dtype=i32, device=dev)
        ax_m = torch.empty(m, dtype=i32, device=dev)
        ctx.tri_boxes(t_m, b_m)
        row = {"prims": m}
        for name, bld in (("lbvh", lr.LBVH), ("ploc", lr.PLOC)):
            ctx.build(b_m, m, nd_m, ax_m, bld, stats=False)   # warm-up
            ts = []
            for _ in range(5):
                s_, e_ = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s_.record(stream)
                ctx.build(b_m, m, nd_m, ax_m, bld, stats=False)
                e_.record(stream)
                ts.append((s_, e_))
            ctx.sync()
            ms = float(np.mean([x.elapsed_time(y) for x, y in ts]))
            row[name + "_ms"] = round(ms, 4)
            row[name + "_mprims_s"] = round(m / (ms * 1e3), 1)
        gbs, frac = rf.roofline(rf.lbvh_bytes(m), row["lbvh_ms"])
        row["lbvh_roofline"] = {"achieved": round(gbs, 1), "unit": "GB/s", "frac": round(frac, 4)}
        live = ctx.build(b_m, m, nd_m, ax_m, lr.PLOC).clusters     # the rounds' live clusters
        gbs, frac = rf.roofline(rf.ploc_bytes(m, live), row["ploc_ms"])
        row["ploc_roofline"] = {"achieved": round(gbs, 1), "unit": "GB/s", "frac": round(frac, 4), "live": live}
        scaling.append(row)
        del t_m, b_m, nd_m, ax_m
    out["build_scaling"] = scaling

    # ---- the CPU baseline: the reference's own CPU path (SAH build, 4-wide walk) ----
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tris_np, rays0, bounce, shadow, nsec, a.cpu_seconds)
    return out


METRIC = "batched Mrays/s + p50/p99 job latency vs plain per-job launch, equal RMSE"
WORKLOAD_BATCHED = ("batched: 64 seeded scenes (triangle soup of 1k-500k triangles, log-uniform; 14 GGX materials "
                    "incl. glass; floor; quad lamp), one job each at a resolution from {256^2, 512^2, 1280x720, "
                    "1920x1080} and spp from {1, 4, 16, 64}, max depth 8, NEE + MIS; all jobs in one engine "
                    "(persistent wavefront kernels, hipGraph of iterations, autotuned launch shape)")


def _align(x):
    return (x + 255) & ~255


def rnd(x, k):
    return None if x is None else round(float(x), k)


def percentile(v, p):
    return float(np.percentile(np.asarray(v, np.float64), p)) if len(v) else None


def build_pool(ctx, lr, plan):
    """one resident scene per job of the plan (scene i: lrs_batch_scene(ntris_i, seed i))"""
    pool = []
    for i, (n, w, h, spp) in enumerate(plan):
        sc = lr.batch_scene(n, i)
        pool.append(lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], 0, sc["eps"], sc["margin"]))
    return pool


def pmc_per_dispatch(counter, args, regex, match, limit_s=300):
    """This script run again in a child process under rocprofv3 --pmc counter (args: its command
    line), the dispatches of kernels matching regex: the mean of the counter over those whose name
    holds match (a list of such names: the sum of their means), and how many. Raises on a
    time-out or no dispatch."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        raise RuntimeError("rocprofv3 not found")
    d = tempfile.mkdtemp(prefix="lr_pmc_", dir="/tmp")
    cmd = [prof, "--pmc", counter, "--kernel-include-regex", regex, "--output-format", "csv", "-d", d, "-o", "run",
           "--", sys.executable, "-u", os.path.abspath(__file__)] + args
    p = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.DEVNULL,
                         stderr=subprocess.DEVNULL, start_new_session=True)
    try:
        p.wait(timeout=limit_s)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        shutil.rmtree(d, ignore_errors=True)
        raise RuntimeError(f"{counter} run passed {limit_s} s")
    names = match if isinstance(match, (list, tuple)) else [match]
    vals = {m: [] for m in names}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for row in csv.DictReader(open(f)):
            for m in names:
                if m in row["Kernel_Name"] and row["Counter_Name"] == counter:
                    vals[m].append(float(row["Counter_Value"]))
    shutil.rmtree(d, ignore_errors=True)
    if not all(vals.values()):
        raise RuntimeError(f"no {counter} dispatches of {[m for m in names if not vals[m]]} (rocprofv3 exit {p.returncode})")
    return sum(sum(v) / len(v) for v in vals.values()), min(len(v) for v in vals.values())


def kernel_traffic(run_args, regex, kname, what):
    """HBM bytes per launch of one kernel, measured for this run: this script run again in child
    processes under rocprofv3 --pmc (run_args: the same work, one step), once with FETCH_SIZE and
    once with WRITE_SIZE, each counter scaled by what it reads for a 512 MiB device copy (two more
    child runs, --copy-probe). Returns (bytes per launch, detail) or (None, the reason)."""
    print(f"bench: HBM traffic of {what} (rocprofv3 --pmc, 4 child runs)", file=sys.stderr, flush=True)
    try:
        probe = 512 << 20
        cf, _ = pmc_per_dispatch("FETCH_SIZE", ["--copy-probe", "512"], "copyBuffer", "copyBuffer")
        cw, _ = pmc_per_dispatch("WRITE_SIZE", ["--copy-probe", "512"], "copyBuffer", "copyBuffer")
        ff, wf = probe / (cf * 1024.0), probe / (cw * 1024.0)
        fk, nf = pmc_per_dispatch("FETCH_SIZE", run_args, regex, kname)
        wk, nw = pmc_per_dispatch("WRITE_SIZE", run_args, regex, kname)
    except Exception as e:   # noqa: BLE001 - reported in the line, the bench goes on
        return None, {"why": str(e)}
    fb, wb = fk * 1024.0 * ff, wk * 1024.0 * wf
    return int(fb + wb), {"kernel": kname, "fetch_bytes": int(fb), "write_bytes": int(wb),
                          "fetch_factor": round(ff, 4), "write_factor": round(wf, 4), "dispatches": [nf, nw],
                          "how": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in child runs of this bench (one step), "
                                 "each scaled by the counter's reading for a 512 MiB device copy"}


def live_traffic(a, tune):
    """HBM bytes per launch of the batched extension kernel (kernel_traffic: the same jobs and
    launch shape, one batch, the extension kernel's dispatches only)."""
    t = tune.as_dict()
    shape = ",".join(str(t[k]) for k in ("block_trace", "stack", "block_shade", "blocks_per_cu", "tile",
                                          "iters_per_graph", "refill", "refill_shadow"))
    kname = (f"k_extend_rf<{t['block_trace']}, {t['stack']}, false>" if t["refill"] > 0
             else f"k_extend<{t['block_trace']}, {t['stack']}, false, false>")
    run_args = ["--tune", shape, "--jobs", str(a.jobs), "--serve-requests", "0", "--no-cpu", "--steps", "1",
                "--warmup", "0", "--skip-plain", "--no-traffic"]
    return kernel_traffic(run_args, "k_extend", kname, "the extension kernel")


def bench_batched(a, torch, lr):
    """The batched workload: one step = every job of the batch through the engine (submit all,
    drain); value = rays traced (closest-hit + shadow, counted on the device) per second of
    wall time over the timed steps. Beside it: the same jobs in the plain per-job form (the
    reference's shape: one job at a time, 1-spp passes, the host waiting for each job), checked
    bit-identical image by image; job latencies; the kernels' times and the roofline of the
    dominant one; serving under Poisson arrivals; the reference CPU path tracer on a sample."""
    steps = a.steps or 2
    warm = 1 if a.warmup is None else a.warmup
    ctx = lr.Context(0)
    plan = lr.batch_plan(a.jobs, 7)
    t0 = time.perf_counter()
    pool = build_pool(ctx, lr, plan)
    build_s = time.perf_counter() - t0
    info = [ps.info() for ps in pool]
    depth = max(d for d, _, _ in info)
    jobs = [lr.make_job(pool[i], lr.batch_camera(1000 + i, w, h), w, h, spp, seed=i, max_depth=8)
            for i, (n, w, h, spp) in enumerate(plan)]
    paths = sum(w * h * spp for (_, w, h, spp) in plan)
    img_bytes = sum(_align(w * h * 24) for (_, w, h, _) in plan)
    big = max(_align(w * h * 24) for (_, w, h, _) in plan)
    eng = lr.Engine(ctx, slots_log2=22, max_tree_depth=depth, image_bytes=img_bytes + big + (64 << 20))
    # the autotuner's sample: the first 24 jobs at 4 spp (about 80M paths, 0.1 s a run), enough for
    # the steady state of the wavefront to outweigh its start and tail, as in the whole batch
    sample = [lr.make_job(pool[i], lr.batch_camera(1000 + i, w, h), w, h, 4, seed=i, max_depth=8)
              for i, (n, w, h, spp) in enumerate(plan[:24])]
    t0 = time.perf_counter()
    if a.tune:
        f = [int(x) for x in a.tune.split(",")]
        eng.tune = dict(zip(["block_trace", "stack", "block_shade", "blocks_per_cu", "tile", "iters_per_graph", "refill",
                                 "refill_shadow"], f))
        tune, tune_mr = eng.tune, 0.0
    else:
        tune, tune_mr = eng.autotune(sample)
    tune_s = time.perf_counter() - t0

    def batch(keep=False):
        s0 = eng.stats()
        t = lr.now_ms()
        ids = [eng.submit(j) for j in jobs]
        assert all(i is not None for i in ids)
        eng.drain()
        t1 = lr.now_ms()
        s1 = eng.stats()
        lat = [eng.job(i).t_done_ms - t for i in ids]
        if not keep:
            for i in ids:
                eng.release(i)
        rays = (s1.rays_ext - s0.rays_ext) + (s1.rays_shadow - s0.rays_shadow)
        return ids, t1 - t, rays, s1.rays_ext - s0.rays_ext, lat, s1.iterations - s0.iterations

    for _ in range(warm):
        batch()
    # ---- the timed region: exactly K batches between full synchronizations ----
    torch.cuda.synchronize()
    rays = 0
    lats, iters = [], 0
    t0 = time.perf_counter()
    for k in range(steps):
        ids, ms, r, rext, lat, it = batch(keep=(k == steps - 1))
        rays += r
        lats += lat
        iters += it
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    value = rays / dt / 1e6
    rays_step = rays // steps

    # ---- plain per-job launches, and the images compared bit for bit ----
    same = 0
    plain_lat, pr0, plain_s = [], eng.stats(), 0.0
    for j, i in zip([] if a.skip_plain else jobs, ids):
        ta = time.perf_counter()
        pid = eng.plain(j)                      # returns when the job is done (the host waited)
        plain_s += time.perf_counter() - ta
        plain_lat.append(plain_s * 1e3)        # one job after another: latency from the batch's start
        x = torch.empty((j.h, j.w, 4), dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        eng.image(i, x)
        eng.image(pid, y)
        same += int(torch.equal(x.view(torch.int32), y.view(torch.int32)))
        eng.release(pid)
        eng.release(i)
    if a.skip_plain:
        for i in ids:
            eng.release(i)
    pr1 = eng.stats()
    plain_rays = (pr1.rays_ext - pr0.rays_ext) + (pr1.rays_shadow - pr0.rays_shadow)
    plain_mrays = plain_rays / plain_s / 1e6 if plain_s > 0 else None
    # the plain form in shortest-first order instead of the batch's (the same per-job times)
    dur = np.diff(np.concatenate([[0.0], plain_lat]))
    order = np.argsort([w * h * spp for (_, w, h, spp) in plan], kind="stable")
    plain_sjf = np.cumsum(dur[order]) if len(dur) == len(plan) else []

    # ---- kernel times (events around each kernel, iterations launched without the graph) and the
    # roofline of the dominant kernel: the extension walk's algorithmic bytes from its own counters
    # (4-wide nodes x 128 B, shapes x 52 B, ray 32 B + hit 16 B per ray) over its time ----
    eng.profiling(True)
    k0 = eng.stats()
    batch()
    k1 = eng.stats()
    eng.profiling(False)
    names = ["regen", "extend", "shade", "shadow"]
    kms = {n: k1.kernel_ms[q] - k0.kernel_ms[q] for q, n in enumerate(names)}
    kl = {n: k1.kernel_launches[q] - k0.kernel_launches[q] for q, n in enumerate(names)}
    eng.counting(True)
    c0 = eng.stats()
    batch()
    c1 = eng.stats()
    eng.counting(False)
    nodes = c1.node_fetches - c0.node_fetches
    shapes = c1.shape_tests - c0.shape_tests
    rext = c1.rays_ext - c0.rays_ext
    alg = rf.walk_bytes(rext, nodes, shapes, rf.NODE_BYTES_WIDE)
    dom = max(kms, key=kms.get)
    achieved = alg / (kms["extend"] * 1e-3) / 1e9
    traffic, traffic_detail = (None, {"why": "--no-traffic"}) if a.no_traffic else live_traffic(a, tune)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(dt * 1e3 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes, cameras and job mix generated on the host; random-init materials)",
        "config": {"workload": WORKLOAD_BATCHED, "jobs": len(jobs), "paths_per_step": paths,
                   "rays_per_step": rays_step, "triangles": sum(n for n, _, _, _ in plan), "max_tree_depth": depth,
                   "slots": 1 << 22, "tune": tune.as_dict(), "tune_mrays_s": round(tune_mr, 1),
                   "tune_s": round(tune_s, 2), "scene_build_s": round(build_s, 2)},
        "batched": {"mrays_s": round(value, 3), "mpaths_s": round(paths * steps / dt / 1e6, 3),
                    "p50_ms": round(percentile(lats, 50), 2), "p99_ms": round(percentile(lats, 99), 2),
                    "iterations_per_step": iters // steps},
        "plain": None if a.skip_plain else {
            "mrays_s": rnd(plain_mrays, 3), "s": rnd(plain_s, 3), "p50_ms": rnd(percentile(plain_lat, 50), 2),
            "p99_ms": rnd(percentile(plain_lat, 99), 2), "rays": plain_rays,
            "shortest_first_p50_ms": rnd(percentile(plain_sjf, 50), 2),
            "shortest_first_p99_ms": rnd(percentile(plain_sjf, 99), 2)},
        "speedup_vs_plain": rnd(value / plain_mrays, 3) if plain_mrays else None,
        "equal_error": {"images_bit_identical": same, "of": 0 if a.skip_plain else len(jobs),
                        "note": "same samples per (job, pixel, index), fixed-point sums: equal RMSE by construction"},
        "kernel_ms_per_launch": {n: round(kms[n] / max(kl[n], 1), 4) for n in names},
        "kernel_ms_total": {n: round(kms[n], 2) for n in names},
        "roofline": {"bound": "hbm", "kernel": "extend", "dominant": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_detail": traffic_detail,
                     "alg_bytes_per_launch": alg // max(kl["extend"], 1), "node_fetches": nodes, "shape_tests": shapes,
                     "rays": rext, "launches": kl["extend"]},
    }
    out.update(walk_counts(c0, c1, kms, kl))
    # shade's roofline: its algorithmic bytes (roofline.shade_bytes: the entries shaded, the paths
    # that went on, the shadow rays made, in the profiled batch) over its event-timed launches
    srext, srsh = k1.rays_ext - k0.rays_ext, k1.rays_shadow - k0.rays_shadow
    salg = rf.shade_bytes(srext, max(srext - paths, 0), srsh)
    sach = salg / (kms["shade"] * 1e-3) / 1e9 if kms["shade"] > 0 else 0.0
    out["shade_roofline"] = {"bound": "hbm", "kernel": "shade", "achieved": round(sach, 2), "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": round(sach / HBM_PEAK_GBS, 4),
                             "alg_bytes_per_launch": salg // max(kl["shade"], 1), "entries": srext,
                             "continued": max(srext - paths, 0), "shadow_rays": srsh}
    if a.serve_requests > 0:
        rates = [float(x) for x in a.serve_rates.split(",") if x]
        # serving answers sooner with shorter graphs (a new job joins at the next graph)
        stune = dict(tune.as_dict(), iters_per_graph=2)
        out["serving"] = bench_serving(ctx, lr, pool[:32], depth, stune, rates, a.serve_requests)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batched(lr, eng, plan, jobs, a.cpu_seconds)
    eng.close()
    for ps in pool:
        ps.close()
    return out


CAUSTICS_W, CAUSTICS_H, CAUSTICS_SPP, CAUSTICS_PHOTONS = 512, 512, 16, 1.0


def bench_caustics(a, torch, lr):
    """The caustics workload: one job of the caustics scene (a glass and a mirror ball in the
    Cornell room), 512x512, 16 spp, one light path a pixel and pass (photons 1: 4.2M light paths
    beside 4.2M camera paths). One step = the job through the engine (submit, drain: its light
    paths at admission, then its camera paths). value = rays traced per second of wall time over
    the timed steps: the path tracer's closest-hit and shadow rays and the light tracer's
    closest-hit rays and rays towards the camera, all counted on the device. The roofline is the
    light kernels' (k_light_first + k_light_rest, a pair a chunk of 2^22 light paths: algorithmic
    bytes from their own counters, 48 B a ray, 64 B a 4-wide node, 52 B a shape, over their time
    from HIP events on the engine's stream; the traffic a pair's)."""
    steps = a.steps or 5
    warm = 1 if a.warmup is None else a.warmup
    W, H, spp, photons = CAUSTICS_W, CAUSTICS_H, CAUSTICS_SPP, CAUSTICS_PHOTONS
    ctx = lr.Context(0)
    sc = lr.caustic_scene(W, H)
    ps = lr.PtScene(ctx, sc["prims"], sc["tris"], sc["mats"], sc["light"], sc["eps"], sc["margin"])
    eng = lr.Engine(ctx, slots_log2=22, max_tree_depth=64, image_bytes=_align(W * H * 24) + (64 << 20))
    if a.tune:
        f = [int(x) for x in a.tune.split(",")]
        eng.tune = dict(zip(["block_trace", "stack", "block_shade", "blocks_per_cu", "tile", "iters_per_graph", "refill",
                             "refill_shadow"], f))
    tune = eng.tune
    job = lr.make_job(ps, sc["cam"], W, H, spp, seed=1, max_depth=8, photons=photons)
    every = lambda st: st.rays_ext + st.rays_shadow + st.light_rays + st.light_shadow_rays

    def step():
        i = eng.submit(job)
        assert i is not None
        eng.drain()
        eng.release(i)

    for _ in range(warm):
        step()
    torch.cuda.synchronize()
    s0 = eng.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s1 = eng.stats()
    d = lambda f: getattr(s1, f) - getattr(s0, f)
    rays = every(s1) - every(s0)
    value = rays / dt / 1e6
    # kernel times (events around each kernel, the light kernel's at admission)
    eng.profiling(True)
    k0 = eng.stats()
    step()
    k1 = eng.stats()
    eng.profiling(False)
    names = ["regen", "extend", "shade", "shadow"]
    kms = {n: k1.kernel_ms[q] - k0.kernel_ms[q] for q, n in enumerate(names)}
    kms["light"] = k1.light_ms - k0.light_ms
    # the light kernel's node records and shapes (its counting form, in a run of its own)
    eng.counting(True)
    c0 = eng.stats()
    step()
    c1 = eng.stats()
    eng.counting(False)
    lrays = (c1.light_rays - c0.light_rays) + (c1.light_shadow_rays - c0.light_shadow_rays)
    lnodes, lshapes = c1.light_node_fetches - c0.light_node_fetches, c1.light_shape_tests - c0.light_shape_tests
    alg = rf.walk_bytes(lrays, lnodes, lshapes, rf.NODE_BYTES_WIDE)
    achieved = alg / (kms["light"] * 1e-3) / 1e9 if kms["light"] > 0 else 0.0
    if a.no_traffic:
        traffic, traffic_detail = None, {"why": "--no-traffic"}
    else:
        t = tune.as_dict()
        shape = ",".join(str(t[k]) for k in ("block_trace", "stack", "block_shade", "blocks_per_cu", "tile",
                                              "iters_per_graph", "refill", "refill_shadow"))
        traffic, traffic_detail = kernel_traffic(["--workload", "caustics", "--tune", shape, "--no-cpu", "--steps", "1",
                                                  "--warmup", "0", "--no-traffic"], "k_light",
                                                 # (COUNT false; the names may go on with the photon form's flag)
                                                 ["k_light_first<256, 16, false", "k_light_rest<256, 16, false"],
                                                 "the light kernels")
    lpaths = d("light_paths") // steps
    chunks = max(1, -(-lpaths // (1 << 22)))
    out = {
        "metric": METRIC + " (caustics workload)",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(dt * 1e3 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (the Cornell room's walls and lamp with a glass and a mirror ball, generated on the host)",
        "config": {"workload": f"caustics: a glass and a mirror ball in the Cornell room, {W}x{H}, {spp} spp, "
                               f"{photons:g} light path a pixel and pass (light tracing + the path tracer with "
                               f"skip_caustics)", "camera_paths_per_step": W * H * spp, "light_paths_per_step": lpaths,
                   "rays_per_step": rays // steps, "tune": tune.as_dict()},
        "light": {"ms": round(kms["light"], 4), "mpaths_s": round(lpaths / kms["light"] / 1e3, 2) if kms["light"] else None,
                  "mrays_s": round(lrays / kms["light"] / 1e3, 2) if kms["light"] else None,
                  "rays_per_path": round(lrays / max(lpaths, 1), 3),
                  "camera_rays_per_path": round((c1.light_shadow_rays - c0.light_shadow_rays) / max(lpaths, 1), 3),
                  "nodes_per_ray": round(lnodes / max(lrays, 1), 2), "shapes_per_ray": round(lshapes / max(lrays, 1), 2)},
        "kernel_ms_total": {n: round(v, 3) for n, v in kms.items()},
        "roofline": {"bound": "hbm", "kernel": "light", "dominant": max(kms, key=kms.get), "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_detail": traffic_detail, "alg_bytes_per_launch": alg // chunks,
                     "node_fetches": lnodes, "shape_tests": lshapes, "rays": lrays, "launches": chunks,
                     "launch": "k_light_first + k_light_rest over one chunk of up to 2^22 light paths"},
    }
    # the photon cache (lrh_photons_*): one frame's light paths into it, a fade, the lookup at W x H
    pc = lr.Photons(ctx, 0.01)
    npc = W * H * spp
    pc.trace(ps, npc, seed=1)
    ctx.sync()
    pt0 = time.perf_counter()
    pc.trace(ps, npc, seed=2)
    ctx.sync()
    pt1 = time.perf_counter()
    pc.fade(0.5)
    ctx.sync()
    pt2 = time.perf_counter()
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    pc.resolve(ps, sc["cam"], W, H, img)   # (its walk stack made on the first call)
    pt3 = time.perf_counter()
    pc.resolve(ps, sc["cam"], W, H, img)
    pt4 = time.perf_counter()
    out["photon_cache"] = {"light_paths": npc, "trace_ms": round((pt1 - pt0) * 1e3, 3),
                           "mpaths_s": round(npc / (pt1 - pt0) / 1e6, 1), "fade_ms": round((pt2 - pt1) * 1e3, 3),
                           "resolve_ms": round((pt4 - pt3) * 1e3, 3),
                           "note": "host wall time around each call (the context's stream synchronised): a trace of "
                                   "W x H x spp light paths into a 2^20-cell table, a fade, the lookup at W x H"}
    pc.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_caustics(sc, W, H, (rays // steps) / spp, a.cpu_seconds)
    eng.close()
    ps.close()
    return out


def cpu_baseline_caustics(sc, W, H, rays_per_pass, seconds):
    """The reference's own CPU path for the same job (oracle/_ref, its float build, OpenMP on
    every host core): trace_image with skip_caustics and light_trace_caustics, one pass (1 spp,
    W x H light paths) at a time, repeated for about `seconds`; the median pass's time, and the
    rays a pass the device traced on the same job (the CPU path does not count rays; both follow
    the same rules)."""
    from oracle import lro
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    threads = min(int(os.environ.get("OMP_NUM_THREADS", cores)), cores)
    if lro.ref(False) is None:
        return {"value": None, "unit": "Mrays/s", "cores": threads, "kind": "reference",
                "sample": "oracle/_ref not built: no CPU baseline", "cpu_model": cpu_model()}
    reps, t_all = [], time.perf_counter()
    while not reps or (time.perf_counter() - t_all < seconds and len(reps) < 1000):
        t0 = time.perf_counter()
        lro.ref_render_lt(sc, W, H, 1, seed=len(reps), trace=True, skip=True, photons=1.0, double=False)
        reps.append(time.perf_counter() - t0)
    med = sorted(reps)[len(reps) // 2]
    return {"value": round(rays_per_pass / med / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "sample": f"{len(reps)} passes of the same job at 1 spp ({W * H} camera paths + {W * H} light paths a pass), "
                      f"{sum(reps):.1f} s; the median pass's time over the device's {rays_per_pass:.0f} rays a pass",
            "pass_s": round(med, 4), "cpu_model": cpu_model()}


def bench_serving(ctx, lr, pool, depth, tune, rates, nreq, target=1e-3):
    """Serving: the pool's scenes resident in a server; for each arrival rate, nreq requests
    (a random scene and camera, a resolution from the batched mix, spp from {4, 16, 64}) arrive
    as a Poisson stream; each is waited for and released by a client thread. Per rate: the rays
    traced per second from the first arrival to the last completion, the completed jobs per
    second, and p50/p99 latency (submission to the last sample). Each request carries the workload's
    target error (1e-3): its first 4 passes decide how many of its spp it takes (the image mean's
    relative standard error at the target; spp_taken: the mean fraction taken)."""
    import threading
    import queue
    budget = sum(ps.info()[2] for ps in pool) + (24 << 30)
    srv = lr.Server(ctx, slots_log2=22, max_tree_depth=depth, budget=budget, tune=tune)
    for ps in pool:
        srv.add_scene(ps)
    res = []
    for ri, lam in enumerate(rates):
        rng = np.random.default_rng(1000 + ri)
        mix = lr.batch_plan(nreq, 200 + ri, spp_lo=4)
        gaps = rng.exponential(1.0 / lam, nreq)
        picks = rng.integers(0, len(pool), nreq)
        q, lat, done_t, frac, errs = queue.Queue(), [], [], [], []

        def collect():
            while True:
                i = q.get()
                if i is None:
                    return
                info = srv.wait(i, 600000.0)
                lat.append(info.t_done_ms - info.t_submit_ms)
                done_t.append(info.t_done_ms)
                frac.append(info.spp_done / info.spp)
                errs.append(info.rel_error)
                srv.release(i)

        th = threading.Thread(target=collect)
        th.start()
        r0 = srv.stats().rays
        t_start = lr.now_ms()
        t_next = time.perf_counter()
        for k in range(nreq):
            t_next += gaps[k]
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            _, w, h, spp = mix[k]
            j = lr.make_job(pool[int(picks[k])], lr.batch_camera(5000 + 1000 * ri + k, w, h), w, h, spp,
                            seed=10000 + k, max_depth=8, target_error=target)
            q.put(srv.submit(j))
        q.put(None)
        th.join()
        r1 = srv.stats().rays
        span = max(done_t) - t_start
        jobs_s = nreq / span * 1e3
        # past saturation (completions well below arrivals) the queue grows for as long as
        # requests come, so latency there measures the stream's length, not the server
        saturated = jobs_s < 0.9 * lam
        res.append({"rate_jobs_s": lam, "requests": nreq, "completions": len(lat), "saturated": bool(saturated),
                    "mrays_s": round((r1 - r0) / span / 1e3, 2), "jobs_s": round(jobs_s, 2),
                    "p50_ms": round(percentile(lat, 50), 1),
                    "p99_ms": round(percentile(lat, 99), 1) if len(lat) >= 100 else None,
                    "span_s": round(span / 1e3, 3), "spp_taken": round(float(np.mean(frac)), 3),
                    "max_rel_error": round(float(max(errs)), 6)})
        print(f"serving: {lam:g} jobs/s: {jobs_s:.1f} jobs/s done, p50 {percentile(lat, 50):.1f} ms"
              f"{' (saturated)' if saturated else ''}", file=sys.stderr, flush=True)
    srv.close()
    sat = [r["rate_jobs_s"] for r in res if not r["saturated"]]
    return {"scenes": len(pool), "rates": res,
            "target_error": target,
            "note": "p99 from >= 100 completions only; saturated: completions < 0.9 x arrivals (latency there grows "
                    "with the stream's length); each job stops at the passes its target error needs (spp_taken)",
            "sustained_rate_jobs_s": max(sat) if sat else None,
            "capacity_jobs_s": max(r["jobs_s"] for r in res)}


def bench_instanced(a, torch, lr):
    """The instanced workload: 2000 dynamic instances (960-triangle spheres, boxes, 64-point
    hulls) bouncing over a 40 x 40 ground, 3000 static props from 20 seeded blobs of 500-20k
    triangles, 14 GGX materials (two glass), a procedural 2048 x 1024 sky with a 5e4 sun. One
    step = one frame: new poses (the moving instances' records only), the top-level BVH rebuilt on the device, a 1280 x 720, 1-spp,
    8-bounce image. value = rays traced per second of frame time. --frames-in-flight F keeps F
    frames rendering at once, each with its own copy of the scene (frame k's top level is built
    while frame k - 1 renders): more rays per iteration, the frames' latency F times one frame's."""
    steps = a.steps or 10
    warm = 2 if a.warmup is None else a.warmup
    ctx = lr.Context(0)
    W, H = 1280, 720
    frames = [lr.instanced_workload(0, f) for f in range(warm + steps)]
    w0 = frames[0]
    F = max(1, a.frames_in_flight)
    sky = lr.sky(2048, 35.0, 60.0, sun_le=(5e4, 4.6e4, 4e4))
    scenes = []
    for _ in range(F):
        s_ = lr.PtScene.instanced(ctx, w0["mtris"], w0["mesh_first"], w0["mats"], w0["instances"])
        s_.set_sky(sky)
        scenes.append(s_)
    ps = scenes[0]
    depth, nshapes, nbytes = ps.info()
    # the top level is rebuilt every frame: its depth moves a little, so the walks get room
    eng = lr.Engine(ctx, slots_log2=22, max_tree_depth=depth + 32, image_bytes=F * _align(W * H * 24) + (16 << 20))
    if a.tune:
        f = [int(x) for x in a.tune.split(",")]
        eng.tune = dict(zip(["block_trace", "stack", "block_shade", "blocks_per_cu", "tile", "iters_per_graph"], f))
    else:
        eng.autotune([lr.make_job(ps, w0["cam"], W, H, 1, seed=99)])
    tune = eng.tune

    builder = {"ploc": lr.PLOC, "lbvh": lr.LBVH, "refit": lr.REFIT}[a.tlas]
    in_flight = {}   # scene copy -> the job rendering it
    # the instances that move (the dynamic ones; the props and the ground keep their poses): a
    # frame sends only their records (lrh_pt_scene_move), then rebuilds the top level
    i0 = w0["instances"]
    moving = np.zeros(len(i0), bool)
    for fr in frames[1:]:
        x = fr["instances"]
        moving |= (x["xform"] != i0["xform"]).any(1) | (x["kind"] != i0["kind"]) | (x["material"] != i0["material"])
    moving = np.nonzero(moving)[0].astype(np.int32)
    if builder == lr.REFIT:   # refits start from (and fall back to) a PLOC top level
        for s_ in scenes:
            s_.move(moving, w0["instances"][moving], builder=lr.PLOC)

    def settle(si):
        """wait for the frame rendering scene copy si, then let it go"""
        i = in_flight.pop(si, None)
        if i is None:
            return
        while eng.job(i).state != lr.JOB_DONE:
            eng.pump(1)
        eng.release(i)

    rigid, phys = None, {"ms": [], "host_ms": [], "contacts": [], "colours": [], "phases": []}
    if a.physics:
        # the dynamic instances as rigid bodies dropped over the ground among the props (fixed):
        # each frame one rigid-body step, and their poses are the frame's moved records
        rigid, moving, scales = lr.instanced_rigid(ctx, w0)
        recs = w0["instances"][moving].copy()

    def frame(k):
        si = k % F
        settle(si)
        if rigid is not None:
            th = time.perf_counter()
            rigid.steps(1.0 / 60, 4, 1)
            recs["xform"] = lr.rigid_xforms(rigid.bodies(0, len(moving)), scales)
            rs = rigid.stats()
            phys["ms"].append(rs.ms)
            phys["contacts"].append(rs.contacts)
            phys["colours"].append(rs.colours)
            phys["phases"].append(list(rs.phase_ms))
            phys["host_ms"].append((time.perf_counter() - th) * 1e3 - rs.ms)
            st = scenes[si].move(moving, recs, builder=builder)
        else:
            st = scenes[si].move(moving, frames[k]["instances"][moving], builder=builder)
        in_flight[si] = eng.submit(lr.make_job(scenes[si], frames[k]["cam"], W, H, 1, seed=k, max_depth=8))
        eng.pump(1)
        return st

    def finish():
        for si in list(in_flight):
            settle(si)

    for k in range(warm):
        frame(k)
    finish()
    torch.cuda.synchronize()
    s0 = eng.stats()
    builds, refits = [], 0
    t0 = time.perf_counter()
    for k in range(warm, warm + steps):
        st = frame(k)
        builds.append(st.ms)
        refits += st.refitted
    finish()
    dt = time.perf_counter() - t0
    s1 = eng.stats()
    rays = (s1.rays_ext - s0.rays_ext) + (s1.rays_shadow - s0.rays_shadow)
    eng.profiling(True)
    k0 = eng.stats()
    frame(0)
    finish()
    k1 = eng.stats()
    eng.profiling(False)
    names = ["regen", "extend", "shade", "shadow"]
    kms = {n: k1.kernel_ms[q] - k0.kernel_ms[q] for q, n in enumerate(names)}
    kl = {n: k1.kernel_launches[q] - k0.kernel_launches[q] for q, n in enumerate(names)}
    eng.counting(True)
    c0 = eng.stats()
    frame(0)
    finish()
    c1 = eng.stats()
    eng.counting(False)
    nodes, shapes, rext = c1.node_fetches - c0.node_fetches, c1.shape_tests - c0.shape_tests, c1.rays_ext - c0.rays_ext
    alg = rf.walk_bytes(rext, nodes, shapes, rf.NODE_BYTES_WIDE)
    achieved = alg / (kms["extend"] * 1e-3) / 1e9
    ninst = len(w0["instances"])
    out = {
        "metric": METRIC + " (instanced workload)",
        "value": round(rays / dt / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(dt * 1e3 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded meshes, poses, materials and sky generated on the host)",
        "config": {"workload": "instanced: 2000 dynamic + 3000 static instances of 23 meshes over a 40x40 ground, "
                               "14 GGX materials, procedural 2048x1024 sky (5e4 sun), TLAS rebuilt per frame, "
                               "1280x720, 1 spp, 8 bounces", "instances": ninst, "mesh_triangles": int(nshapes),
                   "scene_bytes": int(nbytes), "walk_depth": depth, "tune": tune.as_dict(),
                   "frames_in_flight": F},
        "tlas_builder": a.tlas,
        "physics": None if rigid is None else {
            "bodies": len(w0["instances"]) - 1, "moving": int(len(moving)),
            "step_ms": round(float(np.mean(phys["ms"][warm:])), 3),
            "host_ms": round(float(np.mean(phys["host_ms"][warm:])), 3),
            "contacts": int(np.mean(phys["contacts"][warm:])), "colours": int(np.max(phys["colours"][warm:])),
            "phase_ms": dict(zip(["pairs", "contacts", "colours", "solver"],
                                 [round(float(x), 3) for x in np.mean(phys["phases"][warm:], 0)])),
            "note": "one rigid-body step (1/60 s, 4 substeps) a frame on the device, then the moving instances' "
                    "records and the top level (poses read back, as the reference's physics writes Instance.xform)"},
        "records_sent_per_frame": int(len(moving)),
        "tlas_refitted_frames": refits,
        "tlas_build_ms": round(float(np.mean(builds)), 4),
        "tlas_build_mprims_s": round(ninst / (float(np.mean(builds)) * 1e3), 2),
        "kernel_ms_per_launch": {n: round(kms[n] / max(kl[n], 1), 4) for n in names},
        "kernel_ms_total": {n: round(kms[n], 3) for n in names},
        "roofline": {"bound": "hbm", "kernel": "extend", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes_per_launch": alg // max(kl["extend"], 1), "node_fetches": nodes, "shape_tests": shapes,
                     "rays": rext},
        **walk_counts(c0, c1, kms, kl),
        "cpu_baseline": {"value": None, "unit": "Mrays/s", "cores": None, "kind": "reference",
                         "sample": "not run for this workload: the reference's instanced CPU path is not built here"},
    }
    if rigid is not None:
        out["config"]["workload"] += "; dynamic instances moved by the rigid-body step"
        rigid.close()
    eng.close()
    for s_ in scenes:
        s_.close()
    return out


def walk_counts(c0, c1, kms, kl):
    """The counting run's figures for both walks: the shadow walk's roofline (48 B per ray: the
    ray and its light's contribution; 64 B per wide node, 52 B per shape, like the extension
    walk's) over its profiled time, and for each walk the share of a wave's lanes at work in its
    node steps and its shape steps (node records / (64 x the waves' node steps), and so on)."""
    d = lambda f: getattr(c1, f) - getattr(c0, f)
    share = lambda a, b: round(a / (64.0 * b), 4) if b else None
    srays, snodes, sshapes = d("rays_shadow"), d("shadow_node_fetches"), d("shadow_shape_tests")
    salg = rf.walk_bytes(srays, snodes, sshapes, rf.NODE_BYTES_WIDE)
    sach = salg / (kms["shadow"] * 1e-3) / 1e9 if kms["shadow"] > 0 else 0.0
    return {
        "shadow_roofline": {"bound": "hbm", "kernel": "shadow", "achieved": round(sach, 2), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(sach / HBM_PEAK_GBS, 4),
                            "alg_bytes_per_launch": salg // max(kl["shadow"], 1), "node_fetches": snodes,
                            "shape_tests": sshapes, "rays": srays},
        "lane_shares": {"extend_node_steps": share(d("node_fetches"), d("wave_node_steps")),
                        "extend_shape_steps": share(d("shape_tests"), d("wave_shape_steps")),
                        "shadow_node_steps": share(snodes, d("shadow_wave_node_steps")),
                        "shadow_shape_steps": share(sshapes, d("shadow_wave_shape_steps")),
                        "extend_uniform_node_steps": round(d("uniform_node_steps") / max(d("wave_node_steps"), 1), 4),
                        "shadow_uniform_node_steps": round(d("shadow_uniform_node_steps") / max(d("shadow_wave_node_steps"), 1), 4),
                        "extend_node_steps_per_ray": round(d("wave_node_steps") * 64.0 / max(d("rays_ext"), 1), 2),
                        "extend_nodes_per_ray": round(d("node_fetches") / max(d("rays_ext"), 1), 2)},
    }


def cpu_baseline_batched(lr, eng, plan, jobs, seconds):
    """The reference's own CPU path tracer (trace_image of its float build, compiled from its
    sources into oracle/_ref, OpenMP over rows on every host core) on the batched workload's jobs
    at 1 spp, in the batch's order, until about `seconds` of tracing (scene builds excluded). Its
    paths per second times the rays per path the device traced on the same jobs at 1 spp gives
    its Mrays/s (the CPU path does not count rays; both follow the same bounce and shadow rules)."""
    from oracle import lro
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    threads = min(int(os.environ.get("OMP_NUM_THREADS", cores)), cores)
    if lro.ref(False) is None:
        return {"value": None, "unit": "Mrays/s", "cores": threads, "kind": "reference",
                "sample": "oracle/_ref not built: no CPU baseline", "cpu_model": cpu_model()}
    # the first pass takes jobs until a third of the time; then passes over the same jobs
    # until at least three and about `seconds` of tracing: the median pass's rate
    scenes, reps = [], []
    while len(reps) < 3 or (sum(r[0] for r in reps) < seconds * 1e3 and len(reps) < 50):
        ms_pass, paths = 0.0, 0
        for i, (n, w, h, spp) in enumerate(plan):
            if reps and i >= len(scenes):
                break
            if not reps and ms_pass >= seconds * 1e3 / 3:
                break
            if i == len(scenes):
                sc = lr.batch_scene(n, i)
                sc["cam"] = np.asarray(jobs[i].cam[:], np.float32)
                scenes.append(sc)
            _, ms = lro.ref_render(scenes[i], w, h, 1, seed=i, max_depth=8, mis=True, double=False, timed=True)
            ms_pass += ms
            paths += w * h
        reps.append((ms_pass, paths))
    used, cpu_paths = len(scenes), reps[0][1]
    rates = sorted(p / ms for ms, p in reps)
    cpu_ms = sum(ms for ms, _ in reps)
    s0 = eng.stats()
    ids = []
    for j in jobs[:used]:
        one = lr.make_job(j.scene, j.cam[:], j.w, j.h, 1, seed=j.seed, max_depth=8)
        ids.append(eng.submit(one))
    eng.drain()
    s1 = eng.stats()
    for i in ids:
        eng.release(i)
    rpp = ((s1.rays_ext - s0.rays_ext) + (s1.rays_shadow - s0.rays_shadow)) / cpu_paths
    mpaths = rates[len(rates) // 2] / 1e3
    return {"value": round(mpaths * rpp, 4), "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "sample": f"the first {used} jobs of the batch at 1 spp ({cpu_paths} paths a pass), {len(reps)} passes, "
                      f"{cpu_ms / 1e3:.1f} s of trace_image; the median pass's Mpaths/s x the device's {rpp:.3f} "
                      f"rays per path on the same jobs",
            "mpaths_s": round(mpaths, 4), "cpu_model": cpu_model()}


def cpu_baseline(tris_np, rays0, bounce, shadow, nsec, seconds):
    """The reference CPU path on this host (all cores, OpenMP over the rays as its tracer does
    over rows): its own binned-SAH BVH over the same triangles (bvh_build, 4-wide), then every
    ray of one step (primaries, bounces, shadow rays: the same rays the GPU traced), repeated for
    about `seconds`; the median repetition's rate, builds included (one SAH build per
    repetition against the GPU's LBVH + PLOC). Falls back to the oracle's port of the same walk
    where the reference's engine was not built."""
    import torch  # noqa: F401  (the tensors below are torch's)
    from oracle import lro

    n = tris_np.shape[0]
    r0 = np.ascontiguousarray(rays0.cpu().numpy())
    rb = np.ascontiguousarray(bounce[:nsec].cpu().numpy())
    rs = np.ascontiguousarray(shadow[:nsec].cpu().numpy())
    nr = r0.shape[0] + rb.shape[0] + rs.shape[0]
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    threads = min(int(os.environ.get("OMP_NUM_THREADS", cores)), cores)
    R = lro.ref()
    kind = "reference" if R is not None else "port"
    out = np.zeros((r0.shape[0], 4), np.float32)
    reps, t_all = [], time.perf_counter()
    while not reps or (time.perf_counter() - t_all < seconds and len(reps) < 50):
        t0 = time.perf_counter()
        if R is not None:
            h = R.refh_tris_new(lro.ptr(tris_np), n, 1)
            t1 = time.perf_counter()
            for rr, anyh in ((r0, 0), (rb, 0), (rs, 1)):
                R.refh_tris_trace(h, lro.ptr(rr), rr.shape[0], anyh, lro.ptr(out))
            t2 = time.perf_counter()
            R.refh_tris_free(h)
        else:
            sah = lro.Sah(lro.tri_boxes(tris_np), 4, 16, wide=True)
            t1 = time.perf_counter()
            for rr, anyh in ((r0, False), (rb, False), (rs, True)):
                sah.trace(tris_np, rr, any_hit=anyh)
            t2 = time.perf_counter()
        reps.append((t2 - t0, t1 - t0, t2 - t1))
    reps.sort()
    tot, tb, tt = reps[len(reps) // 2]
    return {"value": round(nr / tot / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": kind,
            "sample": f"one whole step's {nr} rays through the CPU's own SAH tree of the same {n} triangles, "
                      f"build included; median of {len(reps)} repetitions",
            "build_s": round(tb, 4), "build_mprims_s": round(n / tb / 1e6, 3), "trace_s": round(tt, 4),
            "trace_only_mrays_s": round(nr / tt / 1e6, 3), "cpu_model": cpu_model()}


def cpu_model():
    """the host CPU's model name (/proc/cpuinfo), for the CPU baseline's record"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
