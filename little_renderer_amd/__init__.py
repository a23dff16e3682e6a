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
REFIT = 2        # an instanced scene's top level: refitted while its cost stays within 1.1 of its build's

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
    _fields_ = [("depth", ctypes.c_int), ("rounds", ctypes.c_int), ("ms", ctypes.c_float),
                ("clusters", ctypes.c_uint64), ("refitted", ctypes.c_int)]


class Tune(ctypes.Structure):
    """LrhTune: the engine's launch shape"""
    _fields_ = [("block_trace", ctypes.c_int), ("stack", ctypes.c_int), ("block_shade", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("tile", ctypes.c_int), ("iters_per_graph", ctypes.c_int),
                ("refill", ctypes.c_int), ("refill_shadow", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Job(ctypes.Structure):
    """LrhJob: one render job (gpu_frame_trace's arguments)"""
    _fields_ = [("scene", ctypes.c_void_p), ("cam", ctypes.c_float * 12), ("w", ctypes.c_int), ("h", ctypes.c_int),
                ("spp", ctypes.c_int), ("s0", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("max_depth", ctypes.c_int),
                ("mis", ctypes.c_int), ("clamp", ctypes.c_float), ("denoise", ctypes.c_int),
                ("target_error", ctypes.c_float), ("photons", ctypes.c_float)]


class JobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("w", ctypes.c_int), ("h", ctypes.c_int), ("spp", ctypes.c_int),
                ("paths", ctypes.c_uint64), ("t_submit_ms", ctypes.c_double), ("t_done_ms", ctypes.c_double),
                ("spp_done", ctypes.c_int), ("rel_error", ctypes.c_float)]


class EngineStats(ctypes.Structure):
    _fields_ = [("rays_ext", ctypes.c_uint64), ("rays_shadow", ctypes.c_uint64), ("iterations", ctypes.c_uint64),
                ("jobs_done", ctypes.c_uint64), ("node_fetches", ctypes.c_uint64), ("shape_tests", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_double * 4), ("kernel_launches", ctypes.c_uint64 * 4),
                ("wave_node_steps", ctypes.c_uint64), ("wave_shape_steps", ctypes.c_uint64),
                ("shadow_node_fetches", ctypes.c_uint64), ("shadow_shape_tests", ctypes.c_uint64),
                ("shadow_wave_node_steps", ctypes.c_uint64), ("shadow_wave_shape_steps", ctypes.c_uint64),
                ("uniform_node_steps", ctypes.c_uint64), ("shadow_uniform_node_steps", ctypes.c_uint64),
                ("light_paths", ctypes.c_uint64), ("light_rays", ctypes.c_uint64), ("light_shadow_rays", ctypes.c_uint64),
                ("light_node_fetches", ctypes.c_uint64), ("light_shape_tests", ctypes.c_uint64),
                ("light_launches", ctypes.c_uint64), ("light_ms", ctypes.c_double)]


class Body(ctypes.Structure):
    """LrhBody: a rigid body (a sphere's radius in half[0]; rot x, y, z, w)"""
    _fields_ = [("shape", ctypes.c_int), ("half", ctypes.c_float * 3), ("pos", ctypes.c_float * 3),
                ("rot", ctypes.c_float * 4), ("vel", ctypes.c_float * 3), ("angvel", ctypes.c_float * 3),
                ("inv_mass", ctypes.c_float), ("inv_inertia", ctypes.c_float * 3), ("restitution", ctypes.c_float),
                ("friction", ctypes.c_float)]


class RigidParams(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float * 3), ("has_room", ctypes.c_int), ("room_lo", ctypes.c_float * 3),
                ("room_hi", ctypes.c_float * 3), ("wall_restitution", ctypes.c_float), ("wall_friction", ctypes.c_float),
                ("iterations", ctypes.c_int), ("linear_damping", ctypes.c_float), ("angular_damping", ctypes.c_float),
                ("warm_start", ctypes.c_int)]


class RigidStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("contacts", ctypes.c_int), ("pairs", ctypes.c_int),
                ("colours", ctypes.c_int), ("ms", ctypes.c_float), ("phase_ms", ctypes.c_float * 4)]


BODY_SPHERE, BODY_BOX = 0, 1


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("admitted", ctypes.c_int64), ("done", ctypes.c_int64),
                ("waiting", ctypes.c_int64), ("rays", ctypes.c_uint64), ("busy_ms", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("budget", ctypes.c_size_t),
                ("in_use", ctypes.c_size_t)]


JOB_RUNNING, JOB_DONE = 1, 2
EAGAIN = -5


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
        "lrh_buffer_alloc": (I, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
        "lrh_buffer_free": (I, [P, P]),
        "lrh_buffer_write": (I, [P, P, P, ctypes.c_size_t]),
        "lrh_buffer_read": (I, [P, P, P, ctypes.c_size_t]),
        "lrh_tri_boxes": (I, [P, P, I, P]),
        "lrh_bounds": (I, [P, P, I, P]),
        "lrh_build": (I, [P, P, I, P, I, U, U, U, P, P, P, ctypes.POINTER(BuildStats)]),
        "lrh_build_noalloc": (I, [P, P, I, P, I, U, U, U, P, P, P, ctypes.POINTER(BuildStats)]),
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
        "lrh_pt_scene_sky": (I, [P, P, I, I, F, F]),
        "lrh_pt_scene_instanced": (I, [P, P, P, I, P, I, P, I, F, I, ctypes.POINTER(P)]),
        "lrh_pt_scene_update": (I, [P, P, I, I, ctypes.POINTER(BuildStats)]),
        "lrh_pt_scene_move": (I, [P, P, P, I, I, ctypes.POINTER(BuildStats)]),
        "lrh_pt_scene_mesh": (I, [P, I, P, I, I, ctypes.POINTER(BuildStats)]),
        "lrh_pt_scene_top": (I, [P, P, P, ctypes.POINTER(ctypes.c_double)]),
        "lrh_pt_scene_trace": (I, [P, P, P, I, I, P]),
        "lrh_pt_scene_info": (I, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_size_t)]),
        "lrh_pt_render": (I, [P, P, P, I, I, I, U, ctypes.c_uint64, I, I, F, P]),
        "lrh_pt_render_job": (I, [P, ctypes.POINTER(Job), P]),
        "lrh_photons_create": (I, [P, F, ctypes.POINTER(P)]),
        "lrh_photons_destroy": (None, [P]),
        "lrh_photons_clear": (I, [P]),
        "lrh_photons_trace": (I, [P, P, U, U, ctypes.c_uint64, I]),
        "lrh_photons_fade": (I, [P, F]),
        "lrh_photons_paths": (ctypes.c_double, [P]),
        "lrh_photons_resolve": (I, [P, P, P, I, I, P]),
        "lrh_photons_read": (I, [P, P]),
        "lrh_bsdf_cases": (I, [P, P, I, P]),
        "lrh_denoise": (I, [P, P, P, P, I, I, I, P]),
        "lrh_engine_guide": (I, [P, ctypes.c_int64, P, P]),
        "lrh_engine_denoise": (I, [P, ctypes.c_int64, I, P]),
        "lrs_batch_scene": (None, [I, ctypes.c_uint64, P, P, P]),
        "lrs_batch_camera": (None, [ctypes.c_uint64, I, I, P]),
        "lrs_batch_plan": (None, [I, ctypes.c_uint64, I, P, P, P, P]),
        "lrs_sky": (None, [I, F, F, P, P]),
        "lrs_sky_tables": (I, [P, I, I, P, P, P, ctypes.POINTER(ctypes.c_double)]),
        "lrs_affine_inverse": (I, [P, P]),
        "lrs_uv_sphere": (I, [I, I, I, P]),
        "lrs_box_mesh": (I, [I, P]),
        "lrs_hull": (I, [I, ctypes.c_uint64, I, P]),
        "lrs_blob": (I, [I, ctypes.c_uint64, I, P]),
        "lrh_now_ms": (ctypes.c_double, []),
        "lrh_engine_create": (I, [P, I, I, ctypes.c_size_t, ctypes.POINTER(P)]),
        "lrh_engine_destroy": (None, [P]),
        "lrh_engine_tune_set": (I, [P, ctypes.POINTER(Tune)]),
        "lrh_engine_tune_get": (I, [P, ctypes.POINTER(Tune)]),
        "lrh_engine_counting": (I, [P, I]),
        "lrh_engine_profiling": (I, [P, I]),
        "lrh_engine_submit": (I, [P, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_int64)]),
        "lrh_engine_pump": (I, [P, I]),
        "lrh_engine_drain": (I, [P]),
        "lrh_engine_plain": (I, [P, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_int64)]),
        "lrh_engine_job": (I, [P, ctypes.c_int64, ctypes.POINTER(JobInfo)]),
        "lrh_engine_image": (I, [P, ctypes.c_int64, P, I]),
        "lrh_engine_release": (I, [P, ctypes.c_int64]),
        "lrh_engine_stats": (I, [P, ctypes.POINTER(EngineStats)]),
        "lrh_engine_autotune": (I, [P, ctypes.POINTER(Job), I, ctypes.POINTER(Tune), ctypes.POINTER(ctypes.c_double)]),
        "lrh_server_create": (I, [P, I, I, ctypes.c_size_t, ctypes.POINTER(Tune), ctypes.POINTER(P)]),
        "lrh_server_add_scene": (I, [P, P]),
        "lrh_server_submit": (I, [P, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_int64)]),
        "lrh_server_wait": (I, [P, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(JobInfo)]),
        "lrh_server_image": (I, [P, ctypes.c_int64, P]),
        "lrh_server_release": (I, [P, ctypes.c_int64]),
        "lrh_server_stats": (I, [P, ctypes.POINTER(ServerStats)]),
        "lrh_server_destroy": (None, [P]),
        "lrh_rigid_create": (I, [P, I, I, ctypes.POINTER(P)]),
        "lrh_rigid_destroy": (None, [P]),
        "lrh_rigid_params_get": (I, [P, ctypes.POINTER(RigidParams)]),
        "lrh_rigid_params_set": (I, [P, ctypes.POINTER(RigidParams)]),
        "lrh_rigid_add_sphere": (I, [P, P, F, F]),
        "lrh_rigid_add_box": (I, [P, P, P, P, F]),
        "lrh_rigid_get": (I, [P, I, I, P]),
        "lrh_rigid_set": (I, [P, I, I, P]),
        "lrh_rigid_steps": (I, [P, F, I, I]),
        "lrh_rigid_stats": (I, [P, ctypes.POINTER(RigidStats)]),
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


def sphere_prim(c, r, material):
    """a sphere in the 96-byte primitive record: centre, radius; q.w 0 (a sphere); u.w material"""
    import numpy as np
    p = np.zeros(24, np.float32)
    p[0:3] = c
    p[3] = r
    p[11] = material
    return p


def caustic_scene(w, h, seed=0):
    """The caustics scene: the Cornell room's five walls and its lamp (lrs_cornell), a clear glass
    ball (material 18) and a mirror ball (19) on the floor: the lamp's light focused through the
    glass and thrown by the mirror lands on the floor and the walls (paths lamp -> glass / mirror
    -> rough -> camera, the ones light tracing adds)"""
    import numpy as np
    sc = cornell(w, h, seed)
    mats = (Material * 20)()
    ctypes.memmove(mats, sc["mats"], ctypes.sizeof(Material) * 18)
    g, m = mats[18], mats[19]
    g.type, g.ior, g.roughness = MAT_GLASS, 1.5, 0.0
    m.type, m.ior, m.roughness = MAT_MIRROR, 1.5, 0.0
    for k in range(3):
        g.albedo[k] = 1.0
        m.albedo[k] = 0.9
    prims = np.stack([sc["prims"][0], sphere_prim((0.62, 0.18, 0.42), 0.18, 18), sphere_prim((0.26, 0.12, 0.30), 0.12, 19)])
    return {"tris": np.ascontiguousarray(sc["tris"][:10]), "prims": prims, "mats": mats, "light": 0, "cam": sc["cam"],
            "eps": 1e-4, "margin": 1e-3}


INST_SPHERE, INST_BOX = -1, -2


def _inst_dtype():
    import numpy as np
    return np.dtype([("xform", np.float32, 12), ("kind", np.int32), ("material", np.int32)])


INSTANCE_DTYPE = _inst_dtype()   # an LrhInstance: xform (object -> world, 3 x 4 rows), kind, material


class PtScene:
    """A scene resident on the device for the path tracer (lrh_pt_scene, lrh_pt_scene_instanced)."""

    @classmethod
    def instanced(cls, ctx, mtris, mesh_first, mats, instances, eps=1e-4, builder=PLOC):
        """lrh_pt_scene_instanced: meshes (triangles in object space, mesh_first: nmeshes + 1
        offsets) and an INSTANCE_DTYPE array of instances"""
        import numpy as np
        self = cls.__new__(cls)
        self.ctx = ctx
        mt = np.ascontiguousarray(mtris, np.float32).reshape(-1, 12)
        mf = np.ascontiguousarray(mesh_first, np.int32)
        inst = np.ascontiguousarray(instances, INSTANCE_DTYPE)
        self._h = ctypes.c_void_p()
        _check(lib().lrh_pt_scene_instanced(ctx._c, _ptr(mt), _ptr(mf), len(mf) - 1, ctypes.cast(mats, ctypes.c_void_p),
                                            len(mats), _ptr(inst), len(inst), eps, builder, ctypes.byref(self._h)))
        self.ninst = len(inst)
        ctx._adopt(self)
        return self

    def __init__(self, ctx, prims, tris, mats, light, eps, margin, builder=PLOC):
        import numpy as np
        self.ctx = ctx
        pr = np.ascontiguousarray(prims, np.float32).reshape(-1, 24) if prims is not None else np.zeros((0, 24), np.float32)
        tr = np.ascontiguousarray(tris, np.float32).reshape(-1, 12) if tris is not None else np.zeros((0, 12), np.float32)
        self._h = ctypes.c_void_p()
        _check(lib().lrh_pt_scene(ctx._c, _ptr(pr), pr.shape[0], _ptr(tr), tr.shape[0], ctypes.cast(mats, ctypes.c_void_p),
                                  len(mats), light, eps, margin, builder, ctypes.byref(self._h)))
        ctx._adopt(self)

    def set_sky(self, rgb, scale=1.0, rot=0.0):
        """lrh_pt_scene_sky: rgb (h, 2h, 3) float32, or None to remove it"""
        import numpy as np
        if rgb is None:
            _check(lib().lrh_pt_scene_sky(self._h, None, 0, 0, 1.0, 0.0))
            return
        m = np.ascontiguousarray(rgb, np.float32)
        _check(lib().lrh_pt_scene_sky(self._h, _ptr(m), m.shape[1], m.shape[0], scale, rot))

    def update(self, instances, builder=LBVH):
        """lrh_pt_scene_update: new poses (an INSTANCE_DTYPE array); the top level's BuildStats"""
        import numpy as np
        inst = np.ascontiguousarray(instances, INSTANCE_DTYPE)
        st = BuildStats()
        _check(lib().lrh_pt_scene_update(self._h, _ptr(inst), len(inst), builder, ctypes.byref(st)))
        return st

    def move(self, idx, instances, builder=LBVH):
        """lrh_pt_scene_move: new poses for the instances idx (each once; an INSTANCE_DTYPE array of
        as many); the top level's BuildStats (zeros when nothing moved)"""
        import numpy as np
        ix = np.ascontiguousarray(idx, np.int32)
        inst = np.ascontiguousarray(instances, INSTANCE_DTYPE)
        if len(ix) != len(inst):
            raise ValueError("one pose per index")
        st = BuildStats()
        _check(lib().lrh_pt_scene_move(self._h, _ptr(ix), _ptr(inst), len(ix), builder, ctypes.byref(st)))
        return st

    def mesh(self, m, tris, builder=PLOC):
        """lrh_pt_scene_mesh: mesh m's new triangles (as many as it has); the mesh build's BuildStats"""
        import numpy as np
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 12)
        st = BuildStats()
        _check(lib().lrh_pt_scene_mesh(self._h, m, _ptr(t), len(t), builder, ctypes.byref(st)))
        return st

    def top(self):
        """lrh_pt_scene_top: (records (2 ninst - 1, 8) as u32, aux (ninst,) as u32, cost)"""
        import numpy as np
        n = self.ninst
        nodes = np.zeros((2 * n - 1, 8), np.uint32)
        aux = np.zeros(n, np.uint32)
        cost = ctypes.c_double()
        _check(lib().lrh_pt_scene_top(self._h, _ptr(nodes), _ptr(aux), ctypes.byref(cost)))
        return nodes, aux, cost.value

    def trace(self, rays, any_hit=False):
        """lrh_pt_scene_trace on host rays (n, 8): hits (n, 4) as a numpy array"""
        import numpy as np
        import torch
        r = torch.from_numpy(np.ascontiguousarray(rays, np.float32).reshape(-1, 8)).cuda()
        h = torch.zeros((r.shape[0], 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _check(lib().lrh_pt_scene_trace(self.ctx._c, self._h, _ptr(r), r.shape[0], 1 if any_hit else 0, _ptr(h)))
        self.ctx.sync()
        return h.cpu().numpy()

    def info(self):
        """(tree depth, shapes, device bytes)"""
        d, n, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        _check(lib().lrh_pt_scene_info(self._h, ctypes.byref(d), ctypes.byref(n), ctypes.byref(b)))
        return d.value, n.value, b.value

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

    def render_job(self, job, accum):
        """lrh_pt_render_job: a job of this scene (make_job: light-traced caustics, a target error)
        added into accum (h, w, 4)"""
        _check(lib().lrh_pt_render_job(self.ctx._c, ctypes.byref(job), _ptr(accum)))


class Photons:
    """The photon cache (lrh_photons_*): light paths' power left at rough vertices in a world-space
    hash grid (cells of `cell`, x4, x16), faded between frames, read where the camera sees a rough
    surface."""
    SLOTS = 1 << 20

    def __init__(self, ctx, cell):
        self.ctx = ctx
        self._p = ctypes.c_void_p()
        _check(lib().lrh_photons_create(ctx._c, cell, ctypes.byref(self._p)))
        ctx._adopt(self)

    def trace(self, scene, npaths, s0=0, seed=0, max_depth=8):
        _check(lib().lrh_photons_trace(self._p, scene._h, npaths, s0, seed, max_depth))

    def fade(self, k):
        _check(lib().lrh_photons_fade(self._p, k))

    def clear(self):
        _check(lib().lrh_photons_clear(self._p))

    def paths(self):
        return lib().lrh_photons_paths(self._p)

    def resolve(self, scene, cam, w, h, out):
        """into a device tensor (h, w, 4): rgb, 1 where a rough surface is seen"""
        import numpy as np
        c = np.ascontiguousarray(cam, np.float32)
        _check(lib().lrh_photons_resolve(self._p, scene._h, _ptr(c), w, h, _ptr(out)))

    def read(self):
        """the table: (SLOTS, 4) uint64"""
        import numpy as np
        t = np.zeros((self.SLOTS, 4), np.uint64)
        _check(lib().lrh_photons_read(self._p, _ptr(t)))
        return t

    def close(self):
        if self._p:
            lib().lrh_photons_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Rigid:
    """Rigid bodies stepped on the device (lrh_rigid_*): spheres and boxes in a room."""

    def __init__(self, ctx, max_bodies=4096, max_contacts=65536):
        self.ctx = ctx
        self._r = ctypes.c_void_p()
        _check(lib().lrh_rigid_create(ctx._c, max_bodies, max_contacts, ctypes.byref(self._r)))
        self._n = 0
        ctx._adopt(self)

    def close(self):
        if self._r:
            lib().lrh_rigid_destroy(self._r)
            self._r = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = RigidParams()
        _check(lib().lrh_rigid_params_get(self._r, ctypes.byref(p)))
        return p

    def set_params(self, **kw):
        """the world's settings changed by name (gravity=(0, -981, 0), room=(lo, hi), ...)"""
        p = self.params
        for k, v in kw.items():
            if k == "room":
                p.has_room = 1
                p.room_lo[:], p.room_hi[:] = list(v[0]), list(v[1])
            elif k == "gravity":
                p.gravity[:] = list(v)
            else:
                setattr(p, k, v)
        _check(lib().lrh_rigid_params_set(self._r, ctypes.byref(p)))

    def add_sphere(self, c, radius, density=1.0):
        i = lib().lrh_rigid_add_sphere(self._r, (ctypes.c_float * 3)(*c), radius, density)
        if i < 0:
            _check(i)
        self._n += 1
        return i

    def add_box(self, c, half, rot=(0.0, 0.0, 0.0, 1.0), density=1.0):
        i = lib().lrh_rigid_add_box(self._r, (ctypes.c_float * 3)(*c), (ctypes.c_float * 3)(*half),
                                    (ctypes.c_float * 4)(*rot), density)
        if i < 0:
            _check(i)
        self._n += 1
        return i

    def bodies(self, first=0, n=None):
        if n is None:
            n = self.count - first
        out = (Body * n)()
        _check(lib().lrh_rigid_get(self._r, first, n, out))
        return out

    def set_bodies(self, first, bodies):
        arr = (Body * len(bodies))(*bodies)
        _check(lib().lrh_rigid_set(self._r, first, len(bodies), arr))

    def steps(self, dt, substeps=4, n=1):
        _check(lib().lrh_rigid_steps(self._r, dt, substeps, n))

    def stats(self):
        st = RigidStats()
        _check(lib().lrh_rigid_stats(self._r, ctypes.byref(st)))
        return st

    @property
    def count(self):
        return self._n


class Context:
    """An lrh context on one device (its own stream unless one is given)."""

    def __init__(self, device=0, stream=None):
        import weakref
        L = lib()
        self._c = ctypes.c_void_p()
        # the scenes, engines and servers made on this context: closed before it (the library's
        # objects use their context's stream and memory until they are freed)
        self._children = weakref.WeakSet()
        _check(L.lrh_create(ctypes.byref(self._c), device, stream))

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if self._c:
            # servers and engines first, then the scenes their jobs read
            for child in sorted(self._children, key=lambda c: isinstance(c, PtScene)):
                child.close()
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

    def denoise(self, accum, moments, guide, w, h, out, passes=5):
        """lrh_denoise on device tensors: accum (h, w, 4), moments (h, w, 2), guide (h, w, 8) -> out (h, w, 4)"""
        _check(lib().lrh_denoise(self._c, _ptr(accum), _ptr(moments), _ptr(guide), w, h, passes, _ptr(out)))

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

    def bsdf_cases(self, cases, out):
        """lrh_bsdf_cases: the device BSDF on (n, 17) cases (device) into (n, 16) (device)"""
        _check(lib().lrh_bsdf_cases(self._c, _ptr(cases), cases.shape[0], _ptr(out)))

    def rays_primary(self, cam, w, h, tmin, tmax, rays):
        _check(lib().lrh_rays_primary(self._c, _ptr(cam), w, h, tmin, tmax, _ptr(rays)))

    def rays_secondary(self, rays, hits, n, tris, light, eps, seed, bounce, shadow, pix, count):
        import numpy as np
        lt = np.ascontiguousarray(light, dtype=np.float32)
        _check(lib().lrh_rays_secondary(self._c, _ptr(rays), _ptr(hits), n, _ptr(tris), _ptr(lt), eps, seed,
                                        _ptr(bounce), _ptr(shadow), _ptr(pix), _ptr(count)))


def now_ms():
    return lib().lrh_now_ms()


def batch_scene(n, seed):
    """A scene of the batched workload (lrs_batch_scene): dict like cornell's"""
    import numpy as np
    tris = np.zeros((n + 2, 12), np.float32)
    prims = np.zeros((1, 24), np.float32)
    mats = (Material * 16)()
    lib().lrs_batch_scene(n, seed, _ptr(tris), _ptr(prims), ctypes.cast(mats, ctypes.c_void_p))
    return {"tris": tris, "prims": prims, "mats": mats, "light": 0, "eps": 1e-4, "margin": 1e-3}


def batch_camera(seed, w, h):
    import numpy as np
    cam = np.zeros(12, np.float32)
    lib().lrs_batch_camera(seed, w, h, _ptr(cam))
    return cam


def batch_plan(njobs, seed, spp_lo=1):
    """The job mix (lrs_batch_plan): list of (ntris, w, h, spp)"""
    import numpy as np
    a = [np.zeros(njobs, np.int32) for _ in range(4)]
    lib().lrs_batch_plan(njobs, seed, spp_lo, *[_ptr(x) for x in a])
    return [tuple(int(x[i]) for x in a) for i in range(njobs)]


def make_job(scene, cam, w, h, spp, s0=0, seed=0, max_depth=8, mis=True, clamp=0.0, denoise=False, target_error=0.0,
             photons=0.0):
    """an LrhJob for a PtScene (denoise: also the denoiser's guide and moments; target_error: stop at
    the pass count whose image mean reaches this relative standard error; photons: light paths per
    pixel and pass traced from the lamp for the caustics)"""
    j = Job()
    j.scene = scene._h.value if isinstance(scene, PtScene) else scene
    for k in range(12):
        j.cam[k] = float(cam[k])
    j.w, j.h, j.spp, j.s0, j.seed = w, h, spp, s0, seed
    j.max_depth, j.mis, j.clamp, j.denoise = max_depth, 1 if mis else 0, clamp, 1 if denoise else 0
    j.target_error = target_error
    j.photons = photons
    return j


class Engine:
    """An engine (lrh_engine_*): many jobs over one pool of path slots."""

    def __init__(self, ctx, slots_log2=22, max_tree_depth=64, image_bytes=1 << 30):
        self.ctx = ctx
        self._e = ctypes.c_void_p()
        _check(lib().lrh_engine_create(ctx._c, slots_log2, max_tree_depth, image_bytes, ctypes.byref(self._e)))
        ctx._adopt(self)

    def close(self):
        if self._e:
            lib().lrh_engine_destroy(self._e)
            self._e = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tune(self):
        t = Tune()
        _check(lib().lrh_engine_tune_get(self._e, ctypes.byref(t)))
        return t

    @tune.setter
    def tune(self, t):
        if isinstance(t, dict):
            x = self.tune   # the keys not given keep their values
            for k, v in t.items():
                setattr(x, k, v)
            t = x
        _check(lib().lrh_engine_tune_set(self._e, ctypes.byref(t)))

    def counting(self, on):
        _check(lib().lrh_engine_counting(self._e, 1 if on else 0))

    def profiling(self, on):
        _check(lib().lrh_engine_profiling(self._e, 1 if on else 0))

    def submit(self, job):
        """the job's id, or None if it does not fit now"""
        i = ctypes.c_int64()
        rc = lib().lrh_engine_submit(self._e, ctypes.byref(job), ctypes.byref(i))
        if rc == EAGAIN:
            return None
        _check(rc)
        return i.value

    def pump(self, graphs=1):
        rc = lib().lrh_engine_pump(self._e, graphs)
        if rc < 0:
            _check(rc)
        return rc

    def drain(self):
        _check(lib().lrh_engine_drain(self._e))

    def plain(self, job):
        i = ctypes.c_int64()
        _check(lib().lrh_engine_plain(self._e, ctypes.byref(job), ctypes.byref(i)))
        return i.value

    def job(self, i):
        info = JobInfo()
        _check(lib().lrh_engine_job(self._e, i, ctypes.byref(info)))
        return info

    def image(self, i, out, add=False):
        """the job's image into a device tensor (h, w, 4): sum of the samples' rgb, count"""
        _check(lib().lrh_engine_image(self._e, i, _ptr(out), 1 if add else 0))

    def guide(self, i, guide, moments):
        """a denoised job's guide (h, w, 8) and moments (h, w, 2) into device tensors"""
        _check(lib().lrh_engine_guide(self._e, i, _ptr(guide), _ptr(moments)))

    def denoise(self, i, out, passes=5):
        """a denoised job's filtered image into a device tensor (h, w, 4): rgb, 1"""
        _check(lib().lrh_engine_denoise(self._e, i, passes, _ptr(out)))

    def release(self, i):
        _check(lib().lrh_engine_release(self._e, i))

    def stats(self):
        st = EngineStats()
        _check(lib().lrh_engine_stats(self._e, ctypes.byref(st)))
        return st

    def autotune(self, jobs):
        arr = (Job * len(jobs))(*jobs)
        best = Tune()
        mr = ctypes.c_double()
        _check(lib().lrh_engine_autotune(self._e, arr, len(jobs), ctypes.byref(best), ctypes.byref(mr)))
        return best, mr.value


class Server:
    """The serving loop (lrh_server_*): resident scenes, requests admitted by memory budget."""

    def __init__(self, ctx, slots_log2=22, max_tree_depth=64, budget=8 << 30, tune=None):
        self.ctx = ctx
        self._s = ctypes.c_void_p()
        t = None
        if tune is not None:
            t = Tune()
            for k, v in (tune.as_dict() if isinstance(tune, Tune) else tune).items():
                setattr(t, k, v)
        _check(lib().lrh_server_create(ctx._c, slots_log2, max_tree_depth, budget,
                                       ctypes.byref(t) if t is not None else None, ctypes.byref(self._s)))
        ctx._adopt(self)

    def close(self):
        if self._s:
            lib().lrh_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene):
        _check(lib().lrh_server_add_scene(self._s, scene._h))

    def submit(self, job):
        i = ctypes.c_int64()
        _check(lib().lrh_server_submit(self._s, ctypes.byref(job), ctypes.byref(i)))
        return i.value

    def wait(self, i, timeout_ms=-1.0):
        """the job's JobInfo when done, None on time-out"""
        info = JobInfo()
        rc = lib().lrh_server_wait(self._s, i, timeout_ms, ctypes.byref(info))
        if rc == EAGAIN:
            return None
        _check(rc)
        return info

    def image(self, i, w, h):
        import numpy as np
        out = np.zeros((h, w, 4), np.float32)
        _check(lib().lrh_server_image(self._s, i, _ptr(out)))
        return out

    def release(self, i):
        _check(lib().lrh_server_release(self._s, i))

    def stats(self):
        st = ServerStats()
        _check(lib().lrh_server_stats(self._s, ctypes.byref(st)))
        return st


def sky(w, sun_elev=35.0, sun_azim=60.0, sun_le=(2600.0, 2400.0, 2100.0)):
    """The procedural sky (lrs_sky): (w/2, w, 3) float32"""
    import numpy as np
    out = np.zeros((w // 2, w, 3), np.float32)
    le = np.ascontiguousarray(sun_le, np.float32)
    lib().lrs_sky(w, sun_elev, sun_azim, _ptr(le), _ptr(out))
    return out


def sky_tables(rgb):
    """lrs_sky_tables: marg (h+1), cond (h*(w+1)), prob (w*h), sum"""
    import numpy as np
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    marg, cond, prob = np.zeros(h + 1, np.float32), np.zeros(h * (w + 1), np.float32), np.zeros(w * h, np.float32)
    sm = ctypes.c_double()
    if lib().lrs_sky_tables(_ptr(rgb), w, h, _ptr(marg), _ptr(cond), _ptr(prob), ctypes.byref(sm)) != 0:
        raise LrhError("lrs_sky_tables failed")
    return marg, cond, prob, sm.value


def affine_inverse(fwd):
    import numpy as np
    f = np.ascontiguousarray(fwd, np.float32).reshape(12)
    inv = np.zeros(12, np.float32)
    if lib().lrs_affine_inverse(_ptr(f), _ptr(inv)) != 0:
        raise LrhError("singular transform")
    return inv


def uv_sphere(slices=30, stacks=17, material=0):
    import numpy as np
    out = np.zeros((2 * slices * (stacks - 1), 12), np.float32)
    n = lib().lrs_uv_sphere(slices, stacks, material, _ptr(out))
    return out[:n]


def box_mesh(material=0):
    import numpy as np
    out = np.zeros((12, 12), np.float32)
    lib().lrs_box_mesh(material, _ptr(out))
    return out


def hull(npts, seed, material=0):
    import numpy as np
    out = np.zeros((npts * npts, 12), np.float32)
    n = lib().lrs_hull(npts, seed, material, _ptr(out))
    if n < 0:
        raise LrhError("lrs_hull: out of memory")
    return out[:n].copy()


def blob(ntris, seed, material=0):
    import numpy as np
    out = np.zeros((ntris + 4 * int(ntris ** 0.5) + 64, 12), np.float32)
    n = lib().lrs_blob(ntris, seed, material, _ptr(out))
    return out[:n].copy()


def pose(rng, scale, at):
    """an object -> world transform (3 x 4): a uniformly random rotation, scale (3,), translation"""
    import numpy as np
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return np.hstack([r @ np.diag(scale), np.asarray(at, np.float64)[:, None]]).astype(np.float32)


CM = 100.0   # the rigid bodies work in centimetres (the reference's physics units), the scenes here in metres


def instanced_rigid(ctx, w, n_dynamic=2000, drop=3.0, seed=0):
    """A rigid world for an instanced_workload scene w: its dynamic instances (1 .. n_dynamic) as
    bodies (the spheres and the 64-point hulls as balls, the boxes as boxes, at their scale),
    dropped from up to `drop` above their places; its static props as fixed balls (scale x 0.8,
    where they stand); the ground as the room's floor with walls at its edges. Units: centimetres
    (the scene's x 100), gravity 981 down, as the reference's physics. Returns (the world, the
    dynamic instances' indices, their scales)."""
    import numpy as np
    inst = w["instances"]
    n = len(inst)
    r = Rigid(ctx, max_bodies=n, max_contacts=max(65536, 16 * n))
    r.set_params(room=((-20 * CM, 0.0, -20 * CM), (20 * CM, 60 * CM, 20 * CM)))
    rng = np.random.default_rng(seed)
    idx, scales, fixed = [], [], []
    for i in range(1, n):
        x = inst[i]["xform"].reshape(3, 4).astype(np.float64)
        s_ = float(np.linalg.norm(x[:, 0]))
        at = x[:, 3] * CM
        kind = int(inst[i]["kind"])
        if i <= n_dynamic:
            q = quat_of(x[:, :3] / s_)
            at = at + np.array([0.0, rng.uniform(0.5, drop) * CM, 0.0])
            if kind == 2:
                r.add_box(tuple(at), (s_ * CM,) * 3, q, 1.0)
            else:
                r.add_sphere(tuple(at), s_ * CM * (1.0 if kind == 1 else 0.95), 1.0)
            idx.append(i)
            scales.append(s_)
        else:
            fixed.append(r.add_sphere(tuple(at), s_ * 0.8 * CM, 1.0))
    if fixed:   # the props do not move: fixed bodies
        b = r.bodies(fixed[0], len(fixed))
        for x in b:
            x.inv_mass = 0.0
            x.inv_inertia[:] = [0.0, 0.0, 0.0]
        r.set_bodies(fixed[0], list(b))
    return r, np.asarray(idx, np.int32), np.asarray(scales, np.float32)


def quat_of(m):
    """the unit quaternion (x, y, z, w) of a rotation matrix (Shepperd's method)"""
    import numpy as np
    t = m[0, 0] + m[1, 1] + m[2, 2]
    if t > 0:
        s_ = np.sqrt(t + 1.0) * 2
        q = ((m[2, 1] - m[1, 2]) / s_, (m[0, 2] - m[2, 0]) / s_, (m[1, 0] - m[0, 1]) / s_, 0.25 * s_)
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s_ = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        q = (0.25 * s_, (m[0, 1] + m[1, 0]) / s_, (m[0, 2] + m[2, 0]) / s_, (m[2, 1] - m[1, 2]) / s_)
    elif m[1, 1] > m[2, 2]:
        s_ = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        q = ((m[0, 1] + m[1, 0]) / s_, 0.25 * s_, (m[1, 2] + m[2, 1]) / s_, (m[0, 2] - m[2, 0]) / s_)
    else:
        s_ = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        q = ((m[0, 2] + m[2, 0]) / s_, (m[1, 2] + m[2, 1]) / s_, 0.25 * s_, (m[1, 0] - m[0, 1]) / s_)
    q = np.asarray(q, np.float64)
    return tuple(q / np.linalg.norm(q))


def rigid_xforms(bodies, scales):
    """the bodies' poses as instance transforms (3 x 4 rows, the scene's units): rotation x scale,
    position / 100"""
    import numpy as np
    n = len(bodies)
    arr = np.ctypeslib.as_array(bodies).view(np.float32).reshape(n, -1)   # Body: 23 words, the shape first
    q = arr[:, 7:11].astype(np.float64)
    p = arr[:, 4:7].astype(np.float64) / CM
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    r = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1).reshape(n, 3, 3)
    r *= np.asarray(scales, np.float64)[:, None, None]
    return np.concatenate([r, p[:, :, None]], 2).reshape(n, 12).astype(np.float32)


def instanced_workload(seed=0, frame=0, n_dynamic=2000, n_static=3000, n_props=20, n_hulls=8):
    """The instanced workload: meshes (ground 40 x 40, a 960-triangle sphere, a box, n_hulls
    64-point convex hulls, n_props seeded blobs of 500-20k triangles, log-uniform), 16
    materials (14 GGX drawn as the batched scenes', two of them glass; the ground; one spare),
    n_dynamic instances (spheres, boxes, hulls; scale 0.2-0.6) bouncing with the frame, and n_static
    props (scale 0.15-0.6), all at uniformly random rotations over the ground.
    Returns dict(mtris, mesh_first, mats, instances, cam for 1280 x 720)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    meshes = []
    ground = np.zeros((2, 12), np.float32)
    q = np.array([-20.0, 0.0, -20.0])
    u, v = np.array([0.0, 0.0, 40.0]), np.array([40.0, 0.0, 0.0])
    for k, (a, b, c) in enumerate(((q, q + u, q + v), (q + u + v, q + v, q + u))):
        ground[k, 0:3], ground[k, 4:7], ground[k, 8:11] = a, b - a, c - a
        ground[k, 3], ground[k, 7] = np.frombuffer(np.uint32(0xffffffff).tobytes(), np.float32)[0], 0.0
        ground[k, 11] = np.frombuffer(np.uint32(14 << 26).tobytes(), np.float32)[0]
    meshes.append(ground)
    meshes.append(uv_sphere(30, 17))
    meshes.append(box_mesh())
    for h in range(n_hulls):
        meshes.append(hull(64, 100 + h))
    for p_ in range(n_props):
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(20000)))))
        meshes.append(blob(n, 200 + p_))
    first = np.cumsum([0] + [len(m) for m in meshes]).astype(np.int32)
    mtris = np.concatenate(meshes)
    sc = batch_scene(1000, seed)          # its 14 GGX materials, the floor's diffuse, a spare
    mats = (Material * 16)()
    ctypes.memmove(mats, sc["mats"], ctypes.sizeof(mats))
    mats[15] = mats[14]
    inst = np.zeros(1 + n_dynamic + n_static, INSTANCE_DTYPE)
    inst[0]["xform"] = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    inst[0]["kind"], inst[0]["material"] = 0, 14
    prng = np.random.default_rng(seed + 1)
    for i in range(n_dynamic):
        kind = (1, 2, 3 + i % n_hulls)[i % 3]
        s = prng.uniform(0.2, 0.6)
        x, z = prng.uniform(-19, 19, 2)
        phase, speed = prng.uniform(0, 2 * np.pi), prng.uniform(0.5, 2.0)
        y = s + 2.0 * abs(np.sin(phase + speed * 0.1 * frame))       # bouncing with the frame
        rot = np.random.default_rng([seed, i, frame])
        inst[1 + i]["xform"] = pose(rot, (s, s, s), (x, y, z)).reshape(12)
        inst[1 + i]["kind"], inst[1 + i]["material"] = kind, int(prng.integers(0, 14))
    for i in range(n_static):
        kind = 3 + n_hulls + int(prng.integers(0, n_props))
        s = prng.uniform(0.15, 0.6)
        x, z = prng.uniform(-19, 19, 2)
        inst[1 + n_dynamic + i]["xform"] = pose(prng, (s, s, s), (x, s * 0.8, z)).reshape(12)
        inst[1 + n_dynamic + i]["kind"], inst[1 + n_dynamic + i]["material"] = kind, int(prng.integers(0, 14))
    cam = camera_make((0.0, 14.0, -30.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 1280 / 720)
    return {"mtris": mtris, "mesh_first": first, "mats": mats, "instances": inst, "cam": cam}
