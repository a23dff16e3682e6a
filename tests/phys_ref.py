"""The reference's own CPU rigid bodies (engine/phys/rigid.c, convex.c), built from its sources
by oracle/Makefile's `phys` target into oracle/_ref/libref_phys.so, behind phys_harness.c's flat
entry points: the checker of the device's rigid-body step. None when it was not built (the
reference's sources are not on every machine)."""
import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "_ref", "libref_phys.so")
_lib = None


def lib():
    global _lib
    if _lib is None and os.path.exists(_PATH):
        L = ctypes.CDLL(_PATH)
        P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        for name, res, args in (("ph_create", P, []), ("ph_free", None, [P]),
                                ("ph_params", None, [P, F, F, F, F, F, I, I, F, F]), ("ph_room", None, [P, P, P]),
                                ("ph_sphere", I, [P, P, F, F]), ("ph_box", I, [P, P, P, P, F]),
                                ("ph_set", None, [P, I, P, P, F, F, I]), ("ph_steps", None, [P, F, I, I]),
                                ("ph_get", None, [P, I, P]), ("ph_contacts", I, [P])):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def f(*a):
    return (ctypes.c_float * len(a))(*a)


class World:
    """a RigidWorld of the reference (rigid_init's defaults; sleeping off unless asked)"""

    def __init__(self, gravity=(0.0, -981.0, 0.0), linear_damping=0.0, angular_damping=0.1, iterations=10, sleep=0,
                 wall_restitution=0.4, wall_friction=0.5):
        self.L = lib()
        self.w = self.L.ph_create()
        self.L.ph_params(self.w, *gravity, linear_damping, angular_damping, iterations, sleep, wall_restitution,
                         wall_friction)
        self.n = 0

    def room(self, lo, hi):
        self.L.ph_room(self.w, f(*lo), f(*hi))

    def sphere(self, c, r, density=1.0):
        self.n += 1
        return self.L.ph_sphere(self.w, f(*c), r, density)

    def box(self, c, half, rot=(0.0, 0.0, 0.0, 1.0), density=1.0):
        self.n += 1
        return self.L.ph_box(self.w, f(*c), f(*half), f(*rot), density)

    def set(self, i, vel=(0, 0, 0), angvel=(0, 0, 0), restitution=0.4, friction=0.5, fixed=False):
        self.L.ph_set(self.w, i, f(*vel), f(*angvel), restitution, friction, 1 if fixed else 0)

    def steps(self, dt, substeps, n):
        self.L.ph_steps(self.w, dt, substeps, n)

    def bodies(self):
        """(n, 14): position, rotation x y z w, velocity, angular velocity, sleeping"""
        out = np.zeros((self.n, 14), np.float32)
        for i in range(self.n):
            o = (ctypes.c_float * 14)()
            self.L.ph_get(self.w, i, o)
            out[i] = list(o)
        return out

    def close(self):
        if self.w:
            self.L.ph_free(self.w)
            self.w = None
