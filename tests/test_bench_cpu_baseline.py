"""CPU tests of the benchmark's CPU-baseline bookkeeping for the batched workload: which jobs a
pass traces, how many passes run, and which pass's rate is reported. The CPU path tracer, the
engine and the scene generator are replaced by stand-ins that record what they are asked to do
(the real ones are exercised by the GPU tests and the benchmark itself)."""
import importlib.util
import os
import sys
import types

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class _Stats:
    def __init__(self, rays):
        self.rays_ext, self.rays_shadow = rays, 0


class _Engine:
    """counts 3 rays per path of every job submitted"""
    def __init__(self):
        self.rays = 0

    def stats(self):
        return _Stats(self.rays)

    def submit(self, job):
        self.rays += 3 * job.w * job.h
        return 1

    def drain(self):
        pass

    def release(self, i):
        pass


class _Job:
    def __init__(self, w, h):
        self.w, self.h, self.scene, self.cam, self.seed = w, h, None, [0.0] * 12, 0


@pytest.fixture
def stand_ins(monkeypatch):
    traced = []
    times = {}

    def ref_render(sc, w, h, spp, seed, max_depth, mis, double, timed):
        traced.append((sc["i"], spp, seed))
        return None, times[sc["i"]]

    lro = types.ModuleType("oracle.lro")
    lro.ref = lambda *a: object()
    lro.ref_render = ref_render
    pkg = types.ModuleType("oracle")
    pkg.lro = lro
    monkeypatch.setitem(sys.modules, "oracle", pkg)
    monkeypatch.setitem(sys.modules, "oracle.lro", lro)
    lr = types.SimpleNamespace(batch_scene=lambda n, i: {"i": i},
                               make_job=lambda sc, cam, w, h, spp, seed, max_depth: _Job(w, h))
    return lr, traced, times


def test_passes_repeat_the_first_pass_jobs_and_report_the_median(stand_ins):
    lr, traced, times = stand_ins
    b = _bench()
    plan = [(1000, 1000, 100, 16)] * 40               # 100k paths a job at 1 spp
    jobs = [_Job(1000, 100) for _ in plan]
    for i in range(40):
        times[i] = 700.0                              # ms a job
    out = b.cpu_baseline_batched(lr, _Engine(), plan, jobs, 10.0)
    # the first pass stops once a third of 10 s is traced: 5 jobs of 0.7 s (3.5 s)
    used = 5
    assert [t[0] for t in traced[:used]] == list(range(used))
    # every later pass traces the same jobs, at 1 spp with their seeds, until >= 3 passes and ~10 s
    assert len(traced) % used == 0
    passes = len(traced) // used
    assert passes >= 3 and passes * used * 0.7 >= 10.0 - 0.7 * used
    assert all(t == (i % used, 1, i % used) for i, t in enumerate(traced))
    # the rate: paths of a pass over the pass's time, times the device's rays per path (3)
    assert out["value"] == pytest.approx(3 * (used * 100000) / (used * 700.0) / 1e3, rel=1e-3)
    assert out["cores"] >= 1 and "cpu_model" in out
    assert f"{passes} passes" in out["sample"]


def test_the_median_pass_is_reported_when_passes_differ(stand_ins):
    lr, traced, times = stand_ins
    b = _bench()
    plan = [(1000, 1000, 1000, 1)] * 2                # 1M paths a job
    jobs = [_Job(1000, 1000) for _ in plan]
    seq = iter([4000.0, 4000.0, 1000.0, 1000.0, 2000.0, 2000.0] + [2000.0] * 100)

    def ref_render(sc, w, h, spp, **k):
        traced.append(sc["i"])
        return None, next(seq)

    sys.modules["oracle.lro"].ref_render = ref_render
    out = b.cpu_baseline_batched(lr, _Engine(), plan, jobs, 30.0)
    # both jobs fit a third of 30 s; passes of 8, 2, 4, 4, 4, 4, 4 s over the same 2M paths (until
    # 30 s): the median is a 4 s pass, not the mean
    assert traced[:6] == [0, 1, 0, 1, 0, 1]
    assert out["mpaths_s"] == pytest.approx(2e6 / 4000.0 / 1e3, rel=1e-3)


def test_no_reference_build_means_no_value(stand_ins, monkeypatch):
    lr, traced, times = stand_ins
    b = _bench()
    monkeypatch.setattr(sys.modules["oracle.lro"], "ref", lambda *a: None)
    out = b.cpu_baseline_batched(lr, _Engine(), [(1000, 4, 4, 1)], [_Job(4, 4)], 10.0)
    assert out["value"] is None and traced == []
