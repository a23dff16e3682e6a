#!/usr/bin/env python3
"""bench.py - little_renderer's path-tracing hot path on one MI355X.

Workload (one step, all inputs resident in HBM before the timed region): the BVH-stress scene
of 1,000,000 seeded triangles (centroids uniform in the unit cube, log-normal edges around
0.002): an LBVH build and a PLOC build (radius 16) from the triangles' boxes, then 1920x1080
pinhole primaries traced (closest hit) through the PLOC tree, and for each primary hit one
cosine-weighted diffuse bounce (closest hit) and one shadow ray toward a fixed point light
(any hit). value = rays traced per second over the whole step (builds included), in Mrays/s.

Prints one JSON line (rank 0). Also: each kernel's time from HIP events recorded on the stream
the kernels run on, the roofline of the dominant kernel (its algorithmic bytes from the walk's
own fetch counters / its average time), and the reference's CPU path (its own SAH build and
4-wide walk, compiled from its sources into oracle/_ref) timed on a bounded sample of the same
rays on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline: repeat for about this long")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 3 times (calibrates the PMC byte counters) and exit")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (the path would shard by "
              f"scene or image tile, see DESIGN.md)", file=sys.stderr)
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
    alg_bytes = 48 * nr + 32 * cn + 52 * ct
    achieved = alg_bytes / (kms[dom] * 1e-3) / 1e9

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
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "alg_bytes": alg_bytes, "node_fetches": cn, "tri_tests": ct, "rays": nr},
    }

    # ---- the CPU baseline: the reference's own CPU path (SAH build, 4-wide walk) ----
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tris_np, rays0, bounce, shadow, nsec, a.cpu_seconds)
    print(json.dumps(out), flush=True)


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
            "build_s": round(tb, 4), "trace_s": round(tt, 4), "trace_only_mrays_s": round(nr / tt / 1e6, 3)}


if __name__ == "__main__":
    main()
