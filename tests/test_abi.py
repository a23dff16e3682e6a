"""CPU tests of the C ABI boundary and the host-side workload generators (no GPU calls).

- Every function include/lr_hip.h and include/lr_scene.h declare is exported by the library.
- The batched / serving workload generators (lrs_batch_plan, lrs_batch_scene, lrs_batch_camera)
  produce what their headers promise: counts, ranges, materials, lamp, floor, camera.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import little_renderer_amd as lr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(lr[hs]_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["lr_hip.h", "lr_scene.h"])
def test_library_exports_every_declared_function(header):
    names = declared(header)
    assert len(names) > 5
    L = ctypes.CDLL(lr.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_abi():
    # every declared entry point has a ctypes signature in the package
    L = lr.lib()
    for h in ("lr_hip.h", "lr_scene.h"):
        for n in declared(h):
            assert getattr(L, n).argtypes is not None or n in ("lrh_last_error",), n


def test_batch_plan_ranges_and_spread():
    plan = lr.batch_plan(4000, 1)
    n = np.array([p[0] for p in plan])
    assert n.min() >= 1000 and n.max() <= 500000
    # log-uniform: the log of the counts is uniform over [ln 1000, ln 500000]
    u = (np.log(n) - math.log(1000)) / (math.log(500000) - math.log(1000))
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
    res = {(p[1], p[2]) for p in plan}
    assert res == {(256, 256), (512, 512), (1280, 720), (1920, 1080)}
    assert {p[3] for p in plan} == {1, 4, 16, 64}
    assert {p[3] for p in lr.batch_plan(400, 1, spp_lo=4)} == {4, 16, 64}
    assert lr.batch_plan(64, 7) == lr.batch_plan(64, 7)
    assert lr.batch_plan(64, 7) != lr.batch_plan(64, 8)


def test_batch_scene_contents():
    n = 5000
    sc = lr.batch_scene(n, 3)
    tris, prims, mats = sc["tris"], sc["prims"], sc["mats"]
    assert tris.shape == (n + 2, 12)
    matid = tris[:, 11].view(np.uint32) >> 26
    assert matid[:n].max() == 13 and matid[:n].min() == 0
    assert np.all(matid[n:] == 14)
    # the soup is the stress generator's
    assert np.array_equal(tris[:n].view(np.uint32), lr.gen_soup(n, 3, 14).view(np.uint32))
    # the floor at y = -0.05 over [-2, 3]^2
    fl = tris[n:]
    assert np.allclose(fl[:, 1], -0.05) and np.allclose(fl[:, 5], 0) and np.allclose(fl[:, 9], 0)
    types = [m.type for m in mats]
    assert types[:14].count(lr.MAT_GLASS) == 2 and types[14] == lr.MAT_DIFFUSE and types[15] == lr.MAT_LIGHT
    for m in mats[:14]:
        if m.type == lr.MAT_PBR:
            assert 0.05 <= m.roughness <= 1.0 and m.metallic in (0.0, 1.0)
            assert all(0.2 <= a <= 0.9 for a in m.albedo)
    assert tuple(mats[15].emission) == (15.0, 15.0, 15.0)
    lamp = prims[0]
    assert lamp[7] == 1.0 and lamp[11] == 15.0                  # a quad, material 15
    assert np.allclose(lamp[16:19], [0, -1, 0])                  # facing down
    assert abs(lamp[15] - 0.25) < 1e-6                           # 0.5 x 0.5


def test_batch_camera_looks_at_the_soup():
    for seed in range(20):
        cam = lr.batch_camera(seed, 1280, 720)
        o = cam[:3].astype(np.float64)
        assert abs(np.linalg.norm(o - 0.5) - 2.2) < 1e-5 and o[1] - 0.5 >= 0.15 * 2.2 - 1e-5
        # the image centre's ray points at (0.5, 0.5, 0.5)
        ul, hz, vt = cam[3:6], cam[6:9], cam[9:12]
        c = ul + 0.5 * hz - 0.5 * vt
        d = (c - o) / np.linalg.norm(c - o)
        t = (np.array([0.5, 0.5, 0.5]) - o) / np.linalg.norm(np.array([0.5, 0.5, 0.5]) - o)
        assert np.dot(d, t) > 1 - 1e-6
        assert abs(np.linalg.norm(hz) / np.linalg.norm(vt) - 1280 / 720) < 1e-4
