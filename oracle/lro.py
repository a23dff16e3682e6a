"""lro.py - ctypes + numpy access to the oracle (liblro.so) and to the reference's own CPU
engine built here (oracle/_ref/libref_{f,d}.so, when it was built).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, and only as the checker or the timed CPU baseline; the product never uses it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblro.so")
LIB_D = os.path.join(HERE, "liblro_d.so")
REF_F = os.path.join(HERE, "_ref", "libref_f.so")
REF_D = os.path.join(HERE, "_ref", "libref_d.so")

P, I, U, U64, F, D = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float,
                      ctypes.c_double)


class LroCount(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint64), ("tris", ctypes.c_uint64)]


_L = None
_LD = None
_R = {}


def lib_d():
    """the oracle's path tracer built in double (liblro_d.so): lro_pt_image at the same samples"""
    global _LD
    if _LD is None:
        if not os.path.exists(LIB_D):
            raise RuntimeError(f"{LIB_D} missing: make -C oracle")
        L = ctypes.CDLL(LIB_D)
        L.lro_pt_image.restype = None
        L.lro_pt_image.argtypes = [P, P, I, I, U, I, U64, F, I, I, P]
        _LD = L
    return _LD


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: make -C oracle")
        L = ctypes.CDLL(LIB)
        sig = {
            "lro_rng_u32": (U, [P]),
            "lro_rng_seed": (None, [P, U64, U64]),
            "lro_rng_float": (F, [P]),
            "lro_rng_double": (D, [P]),
            "lro_tri_boxes": (None, [P, I, P]),
            "lro_union": (None, [P, I, P, P]),
            "lro_inv_extent": (None, [P, P, P]),
            "lro_morton": (None, [P, I, P, P, P]),
            "lro_sort": (None, [P, I, P, P]),
            "lro_lbvh": (I, [P, P, P, I, U, U, U, P, P]),
            "lro_lbvh_bottom_up": (I, [P, P, P, I, U, U, U, P, P]),
            "lro_ploc": (I, [P, P, P, I, U, U, U, P, P, P, P]),
            "lro_check_tree": (I, [P, I, U, P, P]),
            "lro_build_tris": (I, [P, I, I, U, U, U, P, P, P]),
            "lro_sah_build": (P, [P, I, I, I, I]),
            "lro_sah_free": (None, [P]),
            "lro_sah_cost": (D, [P]),
            "lro_sah_info": (None, [P, P, P, P, P]),
            "lro_sah_put_nodes": (None, [P, U, U, P, P]),
            "lro_trace_gpu_layout": (None, [P, U, P, P, P, I, I, P, P]),
            "lro_trace_sah": (None, [P, P, P, I, I, P]),
            "lro_trace_brute": (None, [P, I, P, I, I, P, P]),
            "lro_sampler_hash": (U, [U]),
            "lro_sampler_owen": (U, [U, U]),
            "lro_sampler_sobol1": (U, [U]),
            "lro_blue_noise": (None, [P]),
            "lro_sampler_2d": (None, [P, I, I, U, U, I, P]),
            "lro_bsdf_case": (None, [P, I, P, P, P, P, F, F, F, P]),
            "lro_bsdf_cases": (None, [P, I, P, P]),
            "lro_pt_image": (None, [P, P, I, I, U, I, U64, F, I, I, P]),
            "lro_pt_window": (None, [P, P, I, I, I, I, I, I, U, I, U64, F, I, I, P]),
            "lro_light_trace": (None, [P, P, I, I, U, U, I, U64, I, P]),
            "lro_pc_key": (U64, [I, I, I, I, I]),
            "lro_pc_deposit": (None, [P, F, P, P, P]),
            "lro_pc_deposit_many": (None, [P, F, I, P, P, P]),
            "lro_pc_fade": (None, [P, ctypes.c_double]),
            "lro_photons_trace": (None, [P, P, F, U, U, U64, I]),
            "lro_pc_resolve": (None, [P, P, F, ctypes.c_double, P, I, I, P]),
            "lro_env_prepare": (I, [P, P, I, I, F, F]),
            "lro_env_free": (None, [P]),
            "lro_env_eval": (None, [P, P, P]),
            "lro_env_sample": (None, [P, F, F, P, P, P]),
            "lro_env_pdf": (F, [P, P]),
            "lro_scene_trace": (None, [P, P, I, I, P]),
            "lro_find_emitters": (I, [P, P, P, I, P, I, P, P]),
            "lro_guide": (None, [P, P, I, I, P]),
            "lro_pt_image_moments": (None, [P, P, I, I, U, I, U64, F, I, I, P, P, P]),
            "lro_denoise": (None, [P, P, P, I, I, I, P]),
            "lro_f2h": (ctypes.c_uint16, [F]),
            "lro_h2f": (F, [ctypes.c_uint16]),
            "lro_oct_pack": (None, [P, P]),
            "lro_oct_unpack": (None, [P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _L = L
    return _L


def ref(double=False):
    """The reference's own CPU engine (float or double build), or None where it was not built."""
    path = REF_D if double else REF_F
    if path in _R:
        return _R[path]
    if not os.path.exists(path):
        _R[path] = None
        return None
    L = ctypes.CDLL(path)
    sig = {
        "refh_pcg32": (None, [U64, U64, I, P]),
        "refh_rng_real": (None, [U64, U64, I, P]),
        "refh_sampler_2d": (None, [I, P, P, P, P, U, P]),
        "refh_bvh_tris": (I, [P, I, I, I, P, P, P, P]),
        "refh_trace_tris": (None, [P, I, I, P, I, I, P]),
        "refh_tris_new": (P, [P, I, I]),
        "refh_tris_free": (None, [P]),
        "refh_tris_trace": (None, [P, P, I, I, P]),
        "sampler_hash": (U, [U]),
        "sampler_owen": (U, [U, U]),
        "sampler_sobol1": (U, [U]),
        "blue_noise_mask": (P, []),
        "refh_brdf_tables": (None, [P]),
        "refh_bsdf_case": (None, [P, I, P, P, P, F, F, F, P]),
        "refh_render": (None, [P, I, P, I, P, I, P, I, I, I, U64, I, I, P]),
        "refh_render_t": (None, [P, I, P, I, P, I, P, I, I, I, U64, I, I, P, P]),
        "refh_render_lt": (None, [P, I, P, I, P, I, P, I, P, I, I, I, U64, I, I, I, ctypes.c_double, P]),
        "refh_env_new": (P, [P, I, I, F, F]),
        "refh_env_free": (None, [P]),
        "refh_env_tables": (None, [P, P, P, P, P]),
        "refh_env_eval": (None, [P, P, P]),
        "refh_env_sample": (None, [P, F, F, P, P, P]),
        "refh_env_pdf": (F, [P, P]),
        "refh_sky": (None, [I, F, F, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _R[path] = L
    return L


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---- numpy-level helpers ------------------------------------------------------------------
def pcg32(seed, stream, n):
    L = lib()
    st = (ctypes.c_uint64 * 2)()
    L.lro_rng_seed(st, seed, stream)
    return np.array([L.lro_rng_u32(st) for _ in range(n)], dtype=np.uint32)


def tri_boxes(tris):
    tris = np.ascontiguousarray(tris, dtype=np.float32)
    n = tris.shape[0]
    out = np.empty((n, 8), dtype=np.float32)
    lib().lro_tri_boxes(ptr(tris), n, ptr(out))
    return out


def union(boxes):
    lo = np.empty(3, np.float32)
    hi = np.empty(3, np.float32)
    lib().lro_union(ptr(boxes), boxes.shape[0], ptr(lo), ptr(hi))
    return lo, hi


def inv_extent(lo, hi):
    ext = np.empty(3, np.float32)
    lib().lro_inv_extent(ptr(np.ascontiguousarray(lo, np.float32)), ptr(np.ascontiguousarray(hi, np.float32)), ptr(ext))
    return ext


def morton(boxes, lo, ext):
    n = boxes.shape[0]
    out = np.empty(n, np.uint32)
    lib().lro_morton(ptr(boxes), n, ptr(np.ascontiguousarray(lo, np.float32)), ptr(np.ascontiguousarray(ext, np.float32)),
                     ptr(out))
    return out


def sort(codes):
    codes = np.ascontiguousarray(codes, np.uint32)
    n = codes.shape[0]
    sc = np.empty(n, np.uint32)
    si = np.empty(n, np.uint32)
    lib().lro_sort(ptr(codes), n, ptr(sc), ptr(si))
    return sc, si


def lbvh(sc, si, boxes, node0=0, ref0=0, ref_add=0, bottom_up=False):
    n = sc.shape[0]
    nodes = np.zeros((max(2 * n - 1, 1), 8), np.uint32)
    aux = np.zeros(max(n, 1), np.uint32)
    fn = lib().lro_lbvh_bottom_up if bottom_up else lib().lro_lbvh
    depth = fn(ptr(sc), ptr(si), ptr(boxes), n, node0, ref0, ref_add, ptr(nodes), ptr(aux))
    return nodes, aux, depth


def ploc(sc, si, boxes, node0=0, ref0=0, ref_add=0, live=False):
    n = sc.shape[0]
    nodes = np.zeros((max(2 * n - 1, 1), 8), np.uint32)
    aux = np.zeros(max(n, 1), np.uint32)
    depth = ctypes.c_int()
    lv = ctypes.c_uint64()
    rounds = lib().lro_ploc(ptr(sc), ptr(si), ptr(boxes), n, node0, ref0, ref_add, ptr(nodes), ptr(aux),
                            ctypes.byref(depth), ctypes.byref(lv))
    return (nodes, aux, depth.value, rounds, lv.value) if live else (nodes, aux, depth.value, rounds)


def check_tree(nodes, n, node0=0):
    cost = ctypes.c_double()
    depth = ctypes.c_int()
    ok = lib().lro_check_tree(ptr(np.ascontiguousarray(nodes, np.uint32)), n, node0, ctypes.byref(cost),
                              ctypes.byref(depth))
    return bool(ok), cost.value, depth.value


def build_tris(tris, ploc_=False):
    """The whole build from triangles: (nodes, aux, sorted codes, sorted items, depth)."""
    tris = np.ascontiguousarray(tris, np.float32)
    n = tris.shape[0]
    nodes = np.zeros((2 * n - 1, 8), np.uint32)
    aux = np.zeros(n, np.uint32)
    keys = np.zeros(2 * n, np.uint32)
    depth = lib().lro_build_tris(ptr(tris), n, 1 if ploc_ else 0, 0, 0, 0, ptr(nodes), ptr(aux), ptr(keys))
    return nodes, aux, keys[:n].copy(), keys[n:].copy(), depth


def trace_gpu_layout(nodes, root, aux, tris, rays, any_hit=False, counts=False):
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[0]
    hits = np.zeros((n, 4), np.float32)
    c = LroCount()
    lib().lro_trace_gpu_layout(ptr(np.ascontiguousarray(nodes, np.uint32)), root, ptr(np.ascontiguousarray(aux, np.uint32)),
                               ptr(np.ascontiguousarray(tris, np.float32)), ptr(rays), n, 1 if any_hit else 0, ptr(hits),
                               ctypes.byref(c))
    return (hits, (c.nodes, c.tris)) if counts else hits


class Sah:
    def __init__(self, boxes, max_leaf=4, bins=16, wide=True):
        self.h = lib().lro_sah_build(ptr(np.ascontiguousarray(boxes, np.float32)), boxes.shape[0], max_leaf, bins,
                                     1 if wide else 0)

    def __del__(self):
        try:
            lib().lro_sah_free(self.h)
        except Exception:
            pass

    def cost(self):
        return lib().lro_sah_cost(self.h)

    def info(self):
        v = [ctypes.c_int() for _ in range(4)]
        lib().lro_sah_info(self.h, *[ctypes.byref(x) for x in v])
        return dict(nnodes=v[0].value, nrefs=v[1].value, depth=v[2].value, nwide=v[3].value)

    def put_nodes(self, node0=0, ref0=0):
        inf = self.info()
        nodes = np.zeros((inf["nnodes"], 8), np.uint32)
        aux = np.zeros(inf["nrefs"], np.uint32)
        lib().lro_sah_put_nodes(self.h, node0, ref0, ptr(nodes), ptr(aux))
        return nodes, aux, inf["depth"]

    def trace(self, tris, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        lib().lro_trace_sah(self.h, ptr(np.ascontiguousarray(tris, np.float32)), ptr(rays), rays.shape[0],
                            1 if any_hit else 0, ptr(hits))
        return hits


def trace_brute(tris, rays, any_hit=False):
    rays = np.ascontiguousarray(rays, np.float32)
    tris = np.ascontiguousarray(tris, np.float32)
    hits = np.zeros((rays.shape[0], 4), np.float32)
    second = np.zeros(rays.shape[0], np.float32)
    lib().lro_trace_brute(ptr(tris), tris.shape[0], ptr(rays), rays.shape[0], 1 if any_hit else 0, ptr(hits), ptr(second))
    return hits, second


def hit_ids(hits):
    return np.ascontiguousarray(hits).view(np.uint32)[:, 1]


# ---- the path tracer ------------------------------------------------------------------------
class LroMat(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("albedo", ctypes.c_float * 3), ("emission", ctypes.c_float * 3),
                ("ior", ctypes.c_float), ("metallic", ctypes.c_float), ("roughness", ctypes.c_float),
                ("transmission", ctypes.c_float)]


class LroEnvS(ctypes.Structure):
    _fields_ = [("w", ctypes.c_int), ("h", ctypes.c_int), ("rgb", ctypes.c_void_p), ("marg", ctypes.c_void_p),
                ("cond", ctypes.c_void_p), ("prob", ctypes.c_void_p), ("sum", ctypes.c_double), ("scale", ctypes.c_float),
                ("rot", ctypes.c_float), ("rot_c", ctypes.c_float), ("rot_s", ctypes.c_float)]


# an instance (LroInst): world -> object rows, kind (mesh >= 0, -1 unit sphere, -2 unit box), material
INST_DTYPE = np.dtype([("inv", np.float32, 12), ("kind", np.int32), ("material", np.int32)])
INST_SPHERE, INST_BOX = -1, -2


class LroScene(ctypes.Structure):
    _fields_ = [("prims", ctypes.c_void_p), ("nprims", ctypes.c_int), ("tris", ctypes.c_void_p), ("ntris", ctypes.c_int),
                ("mats", ctypes.c_void_p), ("nmats", ctypes.c_int), ("light", ctypes.c_int), ("eps", ctypes.c_float),
                ("margin", ctypes.c_float), ("mask", ctypes.c_void_p), ("tab", ctypes.c_void_p), ("env", ctypes.c_void_p),
                ("mtris", ctypes.c_void_p), ("mesh_first", ctypes.c_void_p), ("nmeshes", ctypes.c_int),
                ("inst", ctypes.c_void_p), ("ninst", ctypes.c_int), ("mesh_box", ctypes.c_void_p),
                ("tnodes", ctypes.c_void_p), ("taux", ctypes.c_void_p), ("emit", ctypes.c_void_p), ("nemit", ctypes.c_int),
                ("inst_emit", ctypes.c_void_p), ("skip_caustics", ctypes.c_int)]


class Env:
    """The oracle's sky (lro_env_prepare): rgb (h, w, 3) float32, its tables"""

    def __init__(self, rgb, scale=1.0, rot=0.0):
        self.rgb = np.ascontiguousarray(rgb, np.float32)
        self.h, self.w = self.rgb.shape[:2]
        self.s = LroEnvS()
        assert lib().lro_env_prepare(ctypes.byref(self.s), ptr(self.rgb), self.w, self.h, scale, rot) == 0

    def tables(self):
        n = (self.h + 1, self.h * (self.w + 1), self.w * self.h)
        out = []
        for src, k in ((self.s.marg, n[0]), (self.s.cond, n[1]), (self.s.prob, n[2])):
            out.append(np.ctypeslib.as_array(ctypes.cast(src, ctypes.POINTER(ctypes.c_float)), shape=(k,)).copy())
        return out + [self.s.sum]

    def eval(self, d):
        o = np.zeros(3, np.float32)
        lib().lro_env_eval(ctypes.byref(self.s), ptr(np.ascontiguousarray(d, np.float32)), ptr(o))
        return o

    def sample(self, u1, u2):
        d, le, pdf = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
        lib().lro_env_sample(ctypes.byref(self.s), u1, u2, ptr(d), ptr(pdf), ptr(le))
        return d, float(pdf[0]), le

    def pdf(self, d):
        return lib().lro_env_pdf(ctypes.byref(self.s), ptr(np.ascontiguousarray(d, np.float32)))

    def __del__(self):
        try:
            lib().lro_env_free(ctypes.byref(self.s))
        except Exception:
            pass


class RefEnv:
    """The reference's own sky (env_from_rgb of its float build), or None if not built"""

    def __init__(self, rgb, scale=1.0, rot=0.0, double=False):
        self.R = ref(double)
        self.rgb = np.ascontiguousarray(rgb, np.float32)
        self.h, self.w = self.rgb.shape[:2]
        self.e = self.R.refh_env_new(ptr(self.rgb), self.w, self.h, scale, rot)

    def tables(self):
        marg = np.zeros(self.h + 1, np.float32)
        cond = np.zeros(self.h * (self.w + 1), np.float32)
        prob = np.zeros(self.w * self.h, np.float32)
        sm = np.zeros(1, np.float64)
        self.R.refh_env_tables(self.e, ptr(marg), ptr(cond), ptr(prob), ptr(sm))
        return marg, cond, prob, float(sm[0])

    def eval(self, d):
        o = np.zeros(3, np.float32)
        self.R.refh_env_eval(self.e, ptr(np.ascontiguousarray(d, np.float32)), ptr(o))
        return o

    def sample(self, u1, u2):
        d, le, pdf = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
        self.R.refh_env_sample(self.e, u1, u2, ptr(d), ptr(pdf), ptr(le))
        return d, float(pdf[0]), le

    def pdf(self, d):
        return self.R.refh_env_pdf(self.e, ptr(np.ascontiguousarray(d, np.float32)))

    def __del__(self):
        try:
            self.R.refh_env_free(self.e)
        except Exception:
            pass


def ref_sky(w, elev, azim):
    """env_procedural_sky of the reference's float build: (w/2, w, 3) float32, or None"""
    R = ref(False)
    if R is None:
        return None
    out = np.zeros((w // 2, w, 3), np.float32)
    R.refh_sky(w, elev, azim, ptr(out))
    return out


def blue_noise():
    m = np.zeros(4096, np.uint16)
    lib().lro_blue_noise(ptr(m))
    return m


def bsdf_cases(cases, tab):
    """lro_bsdf_cases: (n, 17) float32 cases -> (n, 16) float32"""
    c = np.ascontiguousarray(cases, np.float32)
    out = np.zeros((c.shape[0], 16), np.float32)
    lib().lro_bsdf_cases(ptr(c), c.shape[0], ptr(np.ascontiguousarray(tab, np.float32)), ptr(out))
    return out


def mats_of(mats):
    """Materials (any ctypes array of structures laid out as LroMat) as an LroMat array"""
    out = (LroMat * len(mats))()
    ctypes.memmove(out, mats, ctypes.sizeof(out))
    return out


def pt_image(scene, W, H, s0, ns, seed=0, clamp=0.0, max_depth=8, mis=True, tab=None, mask=None, double=False):
    """lro_pt_image: the rgb sums of samples s0 .. s0+ns-1 of every pixel, (H, W, 3) float32.
    scene: dict of prims (n, 24), tris (m, 12), mats, light, cam, eps, margin; tab, mask: the
    BSDF tables and blue-noise ranks (as the device gets them); double: the FP64 build"""
    sc, keep = scene_struct(scene, tab, mask)
    out = np.zeros((H, W, 3), np.float32)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    (lib_d() if double else lib()).lro_pt_image(ctypes.byref(sc), ptr(cam), W, H, s0, ns, seed, clamp, max_depth, 1 if mis else 0, ptr(out))
    del keep
    return out


def light_trace(scene, W, H, nlp, s0, npass, seed=0, max_depth=8, tab=None, mask=None):
    """lro_light_trace: the splats of npass passes of nlp light paths each (passes s0 ..), (H, W, 3)
    int64 in units of 2^-28 (the engine's fixed point)"""
    sc, keep = scene_struct(scene, tab, mask)
    acc = np.zeros((H, W, 3), np.int64)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    lib().lro_light_trace(ctypes.byref(sc), ptr(cam), W, H, nlp, s0, npass, seed, max_depth, ptr(acc))
    del keep
    return acc


PC_SLOTS = 1 << 20


def pc_table():
    """an empty photon-cache table: (2^20, 4) uint64"""
    return np.zeros((PC_SLOTS, 4), np.uint64)


def photons_trace(scene, table, cell, npaths, s0=0, seed=0, max_depth=8, tab=None):
    """lro_photons_trace: npaths light paths' power left in the table (in place)"""
    sc, keep = scene_struct(scene, tab, None)
    lib().lro_photons_trace(ctypes.byref(sc), ptr(table), cell, npaths, s0, seed, max_depth)
    del keep


def pc_deposit(table, cell, p, n, c):
    """lro_pc_deposit: one photon of power c at p on a surface facing n (in place)"""
    f = lambda x: np.ascontiguousarray(x, np.float32)
    lib().lro_pc_deposit(ptr(table), cell, ptr(f(p)), ptr(f(n)), ptr(f(c)))


def pc_deposit_many(table, cell, p, n, c):
    """lro_pc_deposit_many: photons (m, 3) each of positions p, normals n, powers c, in order"""
    f = lambda x: np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    p, n, c = f(p), f(n), f(c)
    lib().lro_pc_deposit_many(ptr(table), cell, p.shape[0], ptr(p), ptr(n), ptr(c))


def pc_fade(table, k):
    lib().lro_pc_fade(ptr(table), k)


def pc_key(x, y, z, axis, level):
    return lib().lro_pc_key(x, y, z, axis, level)


def pc_resolve(scene, table, cell, paths, W, H, tab=None):
    """lro_pc_resolve: (H, W, 4) float32, rgb and 1 where the pixel's centre ray meets a rough surface"""
    sc, keep = scene_struct(scene, tab, None)
    out = np.zeros((H, W, 4), np.float32)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    lib().lro_pc_resolve(ctypes.byref(sc), ptr(table), cell, paths, ptr(cam), W, H, ptr(out))
    del keep
    return out


def pc_cells(table):
    """the held cells of a table: {tag: (r, g, b, photons x 16)}"""
    t = np.asarray(table)
    held = (t[:, 3] >> np.uint64(32)) != 0
    tags = (t[held, 3] >> np.uint64(32)).astype(np.uint64)
    rows = t[held]
    return {int(g): (int(r[0]), int(r[1]), int(r[2]), int(r[3] & np.uint64(0xffffffff))) for g, r in zip(tags, rows)}


def guide(scene, W, H, tab, mask=None):
    """lro_guide: the denoiser's guide (H, W, 8) float32: n.xyz, z | albedo.rgb, kind (tab: the
    BSDF tables, which decide a surface's kind)"""
    if tab is None:
        raise ValueError("lro.guide needs the BSDF tables")
    sc, keep = scene_struct(scene, tab, mask)
    out = np.zeros((H, W, 8), np.float32)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    lib().lro_guide(ctypes.byref(sc), ptr(cam), W, H, ptr(out))
    del keep
    return out


def pt_image_moments(scene, W, H, s0, ns, guide_, seed=0, clamp=0.0, max_depth=8, mis=True, tab=None, mask=None):
    """lro_pt_image_moments: (sums (H, W, 3), moments (H, W, 2)) for the guide guide_"""
    sc, keep = scene_struct(scene, tab, mask)
    out = np.zeros((H, W, 3), np.float32)
    mom = np.zeros((H, W, 2), np.float32)
    g = np.ascontiguousarray(guide_, np.float32)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    lib().lro_pt_image_moments(ctypes.byref(sc), ptr(cam), W, H, s0, ns, seed, clamp, max_depth, 1 if mis else 0, ptr(g),
                               ptr(out), ptr(mom))
    del keep
    return out, mom


def denoise(accum, moments, guide_, passes=5):
    """lro_denoise: accum (H, W, 4), moments (H, W, 2), guide (H, W, 8) -> (H, W, 4)"""
    a = np.ascontiguousarray(accum, np.float32)
    m = np.ascontiguousarray(moments, np.float32)
    g = np.ascontiguousarray(guide_, np.float32)
    H, W = a.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().lro_denoise(ptr(a), ptr(m), ptr(g), W, H, passes, ptr(out))
    return out


def pt_window(scene, W, H, x0, y0, cw, ch, s0, ns, seed=0, clamp=0.0, max_depth=8, mis=True, tab=None, mask=None):
    """lro_pt_window: pt_image's sums for the pixels [x0, x0+cw) x [y0, y0+ch) of a W x H image,
    (ch, cw, 3) float32"""
    sc, keep = scene_struct(scene, tab, mask)
    out = np.zeros((ch, cw, 3), np.float32)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    lib().lro_pt_window(ctypes.byref(sc), ptr(cam), W, H, x0, y0, cw, ch, s0, ns, seed, clamp, max_depth, 1 if mis else 0,
                        ptr(out))
    del keep
    return out


def scene_struct(scene, tab=None, mask=None):
    """an LroScene for a scene dict (prims, tris, mats, light, eps, margin; optional "sky": an Env,
    "meshes": (tris (m, 12), first (nmeshes + 1,)), "instances": an INST_DTYPE array); and the
    arrays it points into, to keep alive"""
    prims = np.ascontiguousarray(scene.get("prims", np.zeros((0, 24), np.float32)), np.float32).reshape(-1, 24)
    tris = np.ascontiguousarray(scene.get("tris", np.zeros((0, 12), np.float32)), np.float32).reshape(-1, 12)
    mats = mats_of(scene["mats"])
    env = scene.get("sky")
    mt, mf, inst, mbox = None, None, None, None
    if scene.get("instances") is not None:
        mt = np.ascontiguousarray(scene["meshes"][0], np.float32)
        mf = np.ascontiguousarray(scene["meshes"][1], np.int32)
        inst = np.ascontiguousarray(scene["instances"], INST_DTYPE)
        if scene.get("cull", True):
            mbox = mesh_boxes(mt, mf)
    tn, ta = None, None
    if scene.get("tree", len(tris) > 4096) and len(tris) > 1:
        # the oracle's own LBVH over the triangles (pinned against brute force), walked instead of it
        tn, ta = tri_tree(tris)
    em, slot, nem = None, None, 0
    if inst is not None and scene.get("fwd") is not None and scene.get("emitters", True):
        # the glowing shapes (lro_find_emitters) from the instances' object -> world rows
        em, slot, nem = find_emitters(scene["fwd"], inst["kind"], inst["material"], mats)
    keep = [prims, tris, mats, tab, mask, env, mt, mf, inst, mbox, tn, ta, em, slot]
    sc = LroScene(ptr(prims), prims.shape[0], ptr(tris), tris.shape[0], ctypes.cast(mats, ctypes.c_void_p), len(mats),
                  scene.get("light", -1), scene["eps"], scene.get("margin", 1e-3), ptr(mask), ptr(tab),
                  ctypes.addressof(env.s) if env is not None else None, ptr(mt), ptr(mf),
                  0 if mf is None else len(mf) - 1, ptr(inst), 0 if inst is None else len(inst), ptr(mbox),
                  ptr(tn), ptr(ta), ptr(em), nem, ptr(slot), 1 if scene.get("skip_caustics") else 0)
    return sc, keep


EMITTER_FLOATS = 18     # LroEmitter: ball (int), c[3], a[9], r, area, le[3]


def find_emitters(fwd, kind, material, mats):
    """lro_find_emitters: (records (16, 18) float32 as LroEmitter, each instance's slot, count)"""
    f = np.ascontiguousarray(fwd, np.float32).reshape(-1, 12)
    k = np.ascontiguousarray(kind, np.int32)
    m = np.ascontiguousarray(material, np.int32)
    out = np.zeros((16, EMITTER_FLOATS), np.float32)
    slot = np.zeros(max(1, len(k)), np.int32)
    n = lib().lro_find_emitters(ptr(f), ptr(k), ptr(m), len(k), ctypes.cast(mats, ctypes.c_void_p), len(mats), ptr(out),
                                ptr(slot))
    return out, slot, n


_TREES = {}


def tri_tree(tris):
    """the oracle's LBVH over triangles (nodes, aux), kept per triangle array"""
    key = (tris.shape[0], hash(tris[:: max(1, tris.shape[0] // 64)].tobytes()), hash(tris[-1].tobytes()))
    if key not in _TREES:
        nodes, aux, _, _, _ = build_tris(tris)
        _TREES.clear()
        _TREES[key] = (np.ascontiguousarray(nodes, np.uint32), np.ascontiguousarray(aux, np.uint32))
    return _TREES[key]


def mesh_boxes(mtris, first):
    """each mesh's object-space box over its triangles' corners q, q + u, q + v: (nmeshes, 6)"""
    t = np.asarray(mtris, np.float32).reshape(-1, 12).astype(np.float64)
    corners = np.stack([t[:, 0:3], t[:, 0:3] + t[:, 4:7], t[:, 0:3] + t[:, 8:11]], 1)
    out = np.zeros((len(first) - 1, 6), np.float32)
    for m in range(len(first) - 1):
        c = corners[first[m]:first[m + 1]].reshape(-1, 3)
        if len(c):
            out[m, 0:3], out[m, 3:6] = c.min(0), c.max(0)
    return out


def scene_trace(scene, rays, any_hit=False):
    """lro_scene_trace: brute-force hits (n, 4): t, shape id (u32), instance (u32), 0"""
    sc, keep = scene_struct(scene)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    out = np.zeros((r.shape[0], 4), np.float32)
    lib().lro_scene_trace(ctypes.byref(sc), ptr(r), r.shape[0], 1 if any_hit else 0, ptr(out))
    del keep
    return out


def ref_render_lt(scene, W, H, spp, seed=0, max_depth=8, trace=True, skip=False, photons=0.0, double=True):
    """The reference's frame_render_rt with light tracing on the same scene (its double build by
    default): trace_image's mean of spp samples (skip: skip_caustics; trace False: none), plus,
    photons > 0, light_trace_caustics of W H spp photons light paths; (H, W, 3) float64, None
    where oracle/_ref was not built. Spheres: the scene's prims whose q.w is 0 (not the lamp)."""
    R = ref(double)
    if R is None:
        return None
    tris = np.ascontiguousarray(scene["tris"], np.float32)
    pr = np.ascontiguousarray(scene["prims"], np.float32).reshape(-1, 24)
    lamp, lamp_mat = None, -1
    if scene["light"] >= 0:
        p = pr[scene["light"]]
        lamp = np.concatenate([p[4:7], p[8:11], p[12:15]]).astype(np.float32)
        lamp_mat = int(p[11])
    sph = [np.array([*p[0:4], p[11]], np.float32) for i, p in enumerate(pr) if i != scene["light"] and p[7] < 0.5]
    sph = np.ascontiguousarray(np.stack(sph) if sph else np.zeros((0, 5), np.float32))
    m = np.zeros((len(scene["mats"]), 12), np.float32)
    for i, x in enumerate(scene["mats"]):
        m[i, 0] = x.type
        m[i, 1:4] = list(x.albedo)
        m[i, 4:8] = [x.ior, x.metallic, x.roughness, x.transmission]
        m[i, 8:11] = list(x.emission)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    px = np.zeros((H, W, 3), np.float64)
    R.refh_render_lt(ptr(tris), tris.shape[0], ptr(lamp), lamp_mat, ptr(sph), sph.shape[0], ptr(m), m.shape[0], ptr(cam),
                     W, H, spp, seed, max_depth, 1 if trace else 0, 1 if skip else 0, photons, ptr(px))
    return px


def ref_render(scene, W, H, spp, seed=0, max_depth=8, mis=True, double=True, timed=False):
    """The reference's own trace_image (its double build by default) on the same scene: the
    mean of spp samples per pixel, (H, W, 3) float64; None where oracle/_ref was not built.
    timed: also the milliseconds trace_image took (its BVH build excluded): (px, ms)"""
    R = ref(double)
    if R is None:
        return None
    tris = np.ascontiguousarray(scene["tris"], np.float32)
    pr = np.ascontiguousarray(scene["prims"], np.float32).reshape(-1, 24)
    lamp = None
    lamp_mat = -1
    if scene["light"] >= 0:
        p = pr[scene["light"]]
        lamp = np.concatenate([p[4:7], p[8:11], p[12:15]]).astype(np.float32)
        lamp_mat = int(p[11])
    m = np.zeros((len(scene["mats"]), 12), np.float32)
    for i, x in enumerate(scene["mats"]):
        m[i, 0] = x.type
        m[i, 1:4] = list(x.albedo)
        m[i, 4:8] = [x.ior, x.metallic, x.roughness, x.transmission]
        m[i, 8:11] = list(x.emission)
    cam = np.ascontiguousarray(scene["cam"], np.float32)
    px = np.zeros((H, W, 3), np.float64)
    ms = np.zeros(1, np.float64)
    R.refh_render_t(ptr(tris), tris.shape[0], ptr(lamp), lamp_mat, ptr(m), m.shape[0], ptr(cam), W, H, spp, seed,
                    max_depth, 1 if mis else 0, ptr(px), ptr(ms))
    return (px, float(ms[0])) if timed else px
