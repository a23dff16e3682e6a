"""little_renderer_amd - little_renderer's path-tracing hot path on AMD MI355X (gfx950).

The product is liblr_hip.so, a C-ABI library (include/lr_hip.h, include/lr_scene.h) of
hand-written HIP kernels. This module binds it with ctypes so Python code (the tests and the
benchmark) can drive it; torch is used only to own device memory and pass its addresses.
There is no CPU fallback: without the built library every call raises.
"""
import ctypes
import os

__all__ = ["lib", "Context", "LrhError", "gen_soup", "camera_make", "LBVH", "PLOC", "LIB_PATH"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblr_hip.so")
LBVH, PLOC = 0, 1

_lib = None


class LrhError(RuntimeError):
    pass


class Material(ctypes.Structure):
    """LrhMaterial (Material, engine/scene/scene.h:31-45, untextured)"""
    _fields_ = [("type", ctypes.c_int), ("albedo", ctypes.c_float * 3), ("emission", ctypes.c_float * 3),
                ("ior", ctypes.c_float), ("metallic", ctypes.c_float), ("roughness", ctypes.c_float),
                ("transmission", ctypes.c_float)]


MAT_DIFFUSE, MAT_MIRROR, MAT_GLASS, MAT_LIGHT, MAT_PBR = 0, 1, 2, 3, 4


class BuildStats(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("rounds", ctypes.c_int), ("ms", ctypes.c_float)]


def lib():
    """The loaded library; raises if it was not built (no fallback of any kind)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LrhError(f"{LIB_PATH} is missing: build it first (make -C little_renderer_amd)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, U, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_float
    sig = {
        "lrh_last_error": (ctypes.c_char_p, []),
        "lrh_create": (I, [ctypes.POINTER(P), I, P]),
        "lrh_destroy": (None, [P]),
        "lrh_sync": (I, [P]),
        "lrh_stream": (P, [P]),
        "lrh_tri_boxes": (I, [P, P, I, P]),
        "lrh_bounds": (I, [P, P, I, P]),
        "lrh_build": (I, [P, P, I, P, I, U, U, U, P, P, P, ctypes.POINTER(BuildStats)]),
        "lrh_trace": (I, [P, P, U, P, P, P, I, I, P, I]),
        "lrh_trace_queue": (I, [P, P, U, P, P, P, I, P, I, P, I]),
        "lrh_trace_count": (I, [P, P, U, P, P, P, I, P, I, P, I, P]),
        "lrh_rays_primary": (I, [P, P, I, I, F, F, P]),
        "lrh_rays_secondary": (I, [P, P, P, I, P, P, F, U, P, P, P, P]),
        "lrs_gen_soup": (None, [I, ctypes.c_uint64, I, P]),
        "lrs_camera_make": (None, [P, P, P, ctypes.c_double, ctypes.c_double, P]),
        "lrs_blue_noise": (None, [P]),
        "lrs_brdf_tables": (None, [P]),
        "lrs_make_quad": (None, [P, P, P, I, P]),
        "lrs_cornell": (None, [ctypes.c_uint64, I, I, P, P, P, P]),
        "lrh_pt_scene": (I, [P, P, I, P, I, P, I, I, F, F, I, ctypes.POINTER(P)]),
        "lrh_pt_scene_free": (None, [P]),
        "lrh_pt_render": (I, [P, P, P, I, I, I, U, ctypes.c_uint64, I, I, F, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise LrhError(f"{lib().lrh_last_error().decode()} (code {rc})")


def _ptr(t):
    """The device (or host) address of a tensor / numpy array, or None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return t.ctypes.data_as(ctypes.c_void_p)


def gen_soup(n, seed, nmat=1):
    """The BVH-stress triangle soup (lrs_gen_soup) as an (n, 12) float32 numpy array."""
    import numpy as np
    out = np.empty((n, 12), dtype=np.float32)
    lib().lrs_gen_soup(n, seed, nmat, _ptr(out))
    return out


def camera_make(frm, at, up, vfov, aspect):
    import numpy as np
    f = np.asarray(frm, dtype=np.float64)
    a = np.asarray(at, dtype=np.float64)
    u = np.asarray(up, dtype=np.float64)
    cam = np.empty(12, dtype=np.float32)
    lib().lrs_camera_make(_ptr(f), _ptr(a), _ptr(u), float(vfov), float(aspect), _ptr(cam))
    return cam


def blue_noise():
    import numpy as np
    m = np.zeros(4096, np.uint16)
    lib().lrs_blue_noise(_ptr(m))
    return m


def brdf_tables():
    import numpy as np
    t = np.zeros(2112, np.float32)
    lib().lrs_brdf_tables(_ptr(t))
    return t


def make_quad(q, u, v, material):
    import numpy as np
    out = np.zeros(24, np.float32)
    f = [np.ascontiguousarray(x, np.float32) for x in (q, u, v)]
    lib().lrs_make_quad(_ptr(f[0]), _ptr(f[1]), _ptr(f[2]), material, _ptr(out))
    return out


def cornell(w, h, seed=0):
    """The Cornell-style workload scene (lrs_cornell): dict of tris (34, 12), prims (1, 24),
    mats (18 Material), light 0, cam (12,)"""
    import numpy as np
    tris = np.zeros((34, 12), np.float32)
    prims = np.zeros((1, 24), np.float32)
    mats = (Material * 18)()
    cam = np.zeros(12, np.float32)
    lib().lrs_cornell(seed, w, h, _ptr(tris), _ptr(prims), ctypes.cast(mats, ctypes.c_void_p), _ptr(cam))
    return {"tris": tris, "prims": prims, "mats": mats, "light": 0, "cam": cam, "eps": 1e-4, "margin": 1e-3}


class PtScene:
    """A scene resident on the device for the path tracer (lrh_pt_scene)."""

    def __init__(self, ctx, prims, tris, mats, light, eps, margin, builder=PLOC):
        import numpy as np
        self.ctx = ctx
        pr = np.ascontiguousarray(prims, np.float32).reshape(-1, 24) if prims is not None else np.zeros((0, 24), np.float32)
        tr = np.ascontiguousarray(tris, np.float32).reshape(-1, 12) if tris is not None else np.zeros((0, 12), np.float32)
        self._h = ctypes.c_void_p()
        _check(lib().lrh_pt_scene(ctx._c, _ptr(pr), pr.shape[0], _ptr(tr), tr.shape[0], ctypes.cast(mats, ctypes.c_void_p),
                                  len(mats), light, eps, margin, builder, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().lrh_pt_scene_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, cam, w, h, spp, accum, s0=0, seed=0, max_depth=8, mis=True, clamp=0.0):
        import numpy as np
        c = np.ascontiguousarray(cam, np.float32)
        _check(lib().lrh_pt_render(self.ctx._c, self._h, _ptr(c), w, h, spp, s0, seed, max_depth, 1 if mis else 0, clamp,
                                   _ptr(accum)))


class Context:
    """An lrh context on one device (its own stream unless one is given)."""

    def __init__(self, device=0, stream=None):
        L = lib()
        self._c = ctypes.c_void_p()
        _check(L.lrh_create(ctypes.byref(self._c), device, stream))

    def close(self):
        if self._c:
            lib().lrh_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().lrh_stream(self._c)

    def sync(self):
        _check(lib().lrh_sync(self._c))

    def tri_boxes(self, tris, boxes):
        _check(lib().lrh_tri_boxes(self._c, _ptr(tris), tris.shape[0], _ptr(boxes)))

    def bounds(self, boxes, out):
        _check(lib().lrh_bounds(self._c, _ptr(boxes), boxes.shape[0], _ptr(out)))

    def build(self, boxes, n, nodes, aux, builder=LBVH, bounds=None, node0=0, ref0=0, ref_add=0,
              sorted_out=None, stats=True):
        import numpy as np
        hb = None
        if bounds is not None:
            hb = np.ascontiguousarray(bounds, dtype=np.float32)
        st = BuildStats() if stats else None
        _check(lib().lrh_build(self._c, _ptr(boxes), n, _ptr(hb), builder, node0, ref0, ref_add, _ptr(nodes),
                               _ptr(aux), _ptr(sorted_out), ctypes.byref(st) if st is not None else None))
        return st

    def trace(self, nodes, root, aux, tris, rays, n, hits, max_depth, any_hit=False):
        _check(lib().lrh_trace(self._c, _ptr(nodes), root, _ptr(aux), _ptr(tris), _ptr(rays), n,
                               1 if any_hit else 0, _ptr(hits), max_depth))

    def trace_queue(self, nodes, root, aux, tris, rays, max_n, count, hits, max_depth, any_hit=False):
        _check(lib().lrh_trace_queue(self._c, _ptr(nodes), root, _ptr(aux), _ptr(tris), _ptr(rays), max_n,
                                     _ptr(count), 1 if any_hit else 0, _ptr(hits), max_depth))

    def trace_count(self, nodes, root, aux, tris, rays, max_n, count, hits, max_depth, counts, any_hit=False):
        """a walk that also adds its node fetches and triangle tests to counts (2 x uint64, device)"""
        _check(lib().lrh_trace_count(self._c, _ptr(nodes), root, _ptr(aux), _ptr(tris), _ptr(rays), max_n,
                                     _ptr(count), 1 if any_hit else 0, _ptr(hits), max_depth, _ptr(counts)))

    def rays_primary(self, cam, w, h, tmin, tmax, rays):
        _check(lib().lrh_rays_primary(self._c, _ptr(cam), w, h, tmin, tmax, _ptr(rays)))

    def rays_secondary(self, rays, hits, n, tris, light, eps, seed, bounce, shadow, pix, count):
        import numpy as np
        lt = np.ascontiguousarray(light, dtype=np.float32)
        _check(lib().lrh_rays_secondary(self._c, _ptr(rays), _ptr(hits), n, _ptr(tris), _ptr(lt), eps, seed,
                                        _ptr(bounce), _ptr(shadow), _ptr(pix), _ptr(count)))
