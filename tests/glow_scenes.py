"""Scenes of glowing shapes for the oracle and the device (a helper of the tests, not collected).

The reference's own check of its emitters (tests/test_scene.c:118-176): a floor of albedo 0.5
and, 50 units above the point the camera looks at, a glowing ball of radius 10 and emission 10;
that point reads 0.5 * 10 * (10 / 50)^2 = 0.2 (a Lambertian floor lit by a ball of radiance L
filling a cone of half angle asin(r / d): E = L pi (r / d)^2, reflected rho / pi * E). The floor
is a flat box instance here (the device's two-level scenes hold instances only), the ball a unit
sphere instance; or, turned and stretched, a glowing box.
"""
import math

import numpy as np

import little_renderer_amd as lr
from oracle import lro


def _mats():
    mats = (lr.Material * 2)()
    mats[0].type = lr.MAT_DIFFUSE
    mats[0].albedo[:] = [0.5, 0.5, 0.5]
    mats[0].ior, mats[0].roughness = 1.5, 1.0
    # material_pbr(black, 0, 1, 1.5, 0) with emission 10 (test_scene.c:123-124)
    mats[1].type = lr.MAT_PBR
    mats[1].albedo[:] = [0.0, 0.0, 0.0]
    mats[1].ior, mats[1].metallic, mats[1].roughness, mats[1].transmission = 1.5, 0.0, 1.0, 0.0
    mats[1].emission[:] = [10.0, 10.0, 10.0]
    return mats


def axes_xform(rot, scale, t):
    """xform_axes (scene.h:371-373): rotation rot (columns: the shape's axes), scaled per axis,
    then moved to t: 12 floats, object -> world rows"""
    m = np.asarray(rot, np.float64) * np.asarray(scale, np.float64)[None, :]
    out = np.zeros((3, 4), np.float64)
    out[:, :3] = m
    out[:, 3] = t
    return out.astype(np.float32).reshape(12)


def axis_angle(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * k + (1 - math.cos(angle)) * (k @ k)


def glow_scene(box=False, W=200, H=200, sky=None):
    """the scene: a dict for the oracle, and (mtris, mesh_first, mats, instances) for the device.
    box: the glowing shape a turned box of half sizes (12, 8, 10) (test_scene.c:156-160)"""
    mats = _mats()
    # one small mesh far away that no instance uses (a two-level scene's mesh list is not empty)
    mtris = np.zeros((1, 12), np.float32)
    mtris[0, 0:3], mtris[0, 4:7], mtris[0, 8:11] = [9e3, 9e3, 9e3], [1, 0, 0], [0, 1, 0]
    first = np.array([0, 1], np.int32)
    fwd = np.zeros((2, 12), np.float32)
    fwd[0] = axes_xform(np.eye(3), (5000, 1, 5000), (0, -1, 0))   # the floor: its top face at y = 0
    if box:
        fwd[1] = axes_xform(axis_angle((1, 2, 0.5), 0.7), (12, 8, 10), (0, 50, 0))
    else:
        fwd[1] = axes_xform(np.eye(3), (10, 10, 10), (0, 50, 0))
    kinds = np.array([lro.INST_BOX, lro.INST_BOX if box else lro.INST_SPHERE], np.int32)
    inst = np.zeros(2, lr.INSTANCE_DTYPE)
    oi = np.zeros(2, lro.INST_DTYPE)
    for k in range(2):
        inst[k]["xform"], inst[k]["kind"], inst[k]["material"] = fwd[k], kinds[k], k
        oi[k]["inv"], oi[k]["kind"], oi[k]["material"] = lr.affine_inverse(fwd[k]), kinds[k], k
    # the floor point below the shape, seen from the side (test_scene.c:129-130), in a narrow view
    cam = lr.camera_make((-200, 20, 0), (0, 0, 0), (0, 1, 0), 0.02, W / H)
    scene = {"prims": np.zeros((0, 24), np.float32), "tris": np.zeros((0, 12), np.float32), "mats": mats, "light": -1,
             "cam": cam, "eps": 2e-3, "margin": 2e-2, "meshes": (mtris, first), "instances": oi, "fwd": fwd}
    if sky is not None:
        scene["sky"] = lro.Env(sky)
    return scene, (mtris, first, mats, inst)
