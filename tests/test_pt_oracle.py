"""CPU tests: pin the oracle's path tracer pieces (oracle/lro_pt.c) and the product's host-side
tables against the reference's own engine (oracle/_ref, where it was built), closed forms and
known answers.
"""
import ctypes

import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro


def need_ref(double=False):
    R = lro.ref(double)
    if R is None:
        pytest.skip("oracle/_ref not built (no /root/reference here)")
    return R


def ref_mask(double=False):
    R = need_ref(double)
    R.blue_noise_mask.restype = ctypes.POINTER(ctypes.c_uint16)
    return np.ctypeslib.as_array(R.blue_noise_mask(), shape=(4096,)).copy()


# ---- the sampler -----------------------------------------------------------------------------
def test_blue_noise_is_a_permutation_and_matches_reference():
    m = lro.blue_noise()
    assert np.array_equal(np.sort(m), np.arange(4096, dtype=np.uint16))
    assert np.array_equal(m, lr.blue_noise())        # the product's own restatement
    R = lro.ref()
    if R is not None:
        # the reference's float build (its GPU takes this mask: trace_gpu.c:524-528)
        assert np.array_equal(m, ref_mask(False))


def test_sampler_pieces_closed_forms():
    L = lro.lib()
    rng = np.random.default_rng(0)
    for x in rng.integers(0, 2**32, 200, dtype=np.uint64):
        x = int(x)
        # sobol1: the second Sobol dimension by its direction numbers v1 = 2^31, v_k+1 = v_k ^ (v_k >> 1)
        v, want = 1 << 31, 0
        for b in range(32):
            if (x >> b) & 1:
                want ^= v
            v ^= v >> 1
        assert L.lro_sampler_sobol1(x) == want
        # lowbias32
        h = x
        h ^= h >> 16; h = (h * 0x7feb352d) & 0xffffffff
        h ^= h >> 15; h = (h * 0x846ca68b) & 0xffffffff
        h ^= h >> 16
        assert L.lro_sampler_hash(x) == h


def test_owen_scrambling_keeps_stratification():
    # the first 2^k points of the scrambled (0,2)-sequence put one point in each of 2^k equal
    # squares (k even: a sqrt grid), whatever the scramble
    mask = lro.blue_noise()
    for k in (4, 6, 8):
        n = 1 << k
        g = 1 << (k // 2)
        for seed in (0, 12345):
            pts = np.zeros((n, 2), np.float32)
            for i in range(n):
                o = np.zeros(2, np.float32)
                lro.lib().lro_sampler_2d(lro.ptr(np.zeros(4096, np.uint16)), 0, 0, i, seed, 1, lro.ptr(o))
                pts[i] = o
            cells = set((int(a * g), int(b * g)) for a, b in pts - 0.5 / 4096)
            assert len(cells) == n
    del mask


def test_sampler_matches_reference():
    R = need_ref(False)
    mask = lro.blue_noise()
    rng = np.random.default_rng(1)
    n = 5000
    x = rng.integers(0, 2000, n).astype(np.int32)
    y = rng.integers(0, 2000, n).astype(np.int32)
    idx = rng.integers(0, 1 << 20, n).astype(np.uint32)
    dim = rng.integers(0, 13, n).astype(np.int32)
    want = np.zeros(2 * n, np.float64)
    R.refh_sampler_2d(n, lro.ptr(x), lro.ptr(y), lro.ptr(idx), lro.ptr(dim), 0x9e3779b9, lro.ptr(want))
    got = np.zeros(2 * n, np.float32)
    o = np.zeros(2, np.float32)
    for i in range(n):
        lro.lib().lro_sampler_2d(lro.ptr(mask), int(x[i]), int(y[i]), int(idx[i]), 0x9e3779b9, int(dim[i]), lro.ptr(o))
        got[2 * i:2 * i + 2] = o
    assert np.array_equal(got.astype(np.float64), want)


# ---- the BSDF --------------------------------------------------------------------------------
def test_brdf_tables_match_reference():
    R = need_ref(False)
    ref = np.zeros(2112, np.float32)
    R.refh_brdf_tables(lro.ptr(ref))
    assert np.array_equal(lr.brdf_tables(), ref)


def test_brdf_tables_known_values():
    # a perfectly smooth white lobe keeps everything: E(mu, r=0) -> 1 away from grazing; B at
    # r = 0 and normal incidence is 0 ((1 - v.h)^5 = 0); rougher lobes lose energy
    t = lr.brdf_tables()
    E = t[:1024].reshape(32, 32)
    B = t[1024:2048].reshape(32, 32)
    assert np.all(np.abs(E[0, 4:] - 1) < 1e-4)
    assert B[0, 31] < 1e-5
    assert E[31, 31] < E[8, 31] < E[0, 31] + 1e-6


def random_cases(n, seed):
    rng = np.random.default_rng(seed)

    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    nrm = unit(n)
    v = unit(n)
    v[np.sum(v * nrm, 1) < 0] *= -1                    # v on the normal's side
    l = unit(n)
    types = rng.integers(0, 5, n)
    types[types == 3] = 4                              # no lamps
    m = np.zeros((n, 9), np.float32)
    m[:, 0] = types
    m[:, 1:4] = rng.uniform(0.05, 1.0, (n, 3))
    m[:, 4] = rng.uniform(1.1, 2.0, n)
    m[:, 5] = rng.uniform(0, 1, n) * (rng.uniform(0, 1, n) < 0.6)
    m[:, 6] = np.where(rng.uniform(0, 1, n) < 0.15, 0.0, rng.uniform(0, 1, n))
    m[:, 7] = rng.uniform(0, 1, n) * (rng.uniform(0, 1, n) < 0.4)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    outside = rng.integers(0, 2, n)
    return nrm.astype(np.float32), v.astype(np.float32), l.astype(np.float32), m, u, outside


def mats_row(x):
    """a Material as the 9 floats of the BSDF cases: type, albedo, ior, metallic, roughness, transmission"""
    return np.array([x.type, *x.albedo, x.ior, x.metallic, x.roughness, x.transmission, 0], np.float32)


def oracle_case(m, outside, tab, n, v, l, u):
    mat = lro.LroMat()
    mat.type = int(m[0])
    mat.albedo[:] = [float(x) for x in m[1:4]]
    mat.ior, mat.metallic, mat.roughness, mat.transmission = [float(x) for x in m[4:8]]
    out = np.zeros(16, np.float32)
    lro.lib().lro_bsdf_case(ctypes.byref(mat), int(outside), lro.ptr(tab), lro.ptr(n), lro.ptr(v), lro.ptr(l),
                            float(u[0]), float(u[1]), float(u[2]), lro.ptr(out))
    return out


def test_bsdf_matches_reference_float_build():
    R = need_ref(False)
    tab = lr.brdf_tables()
    n, v, l, m, u, outside = random_cases(4000, 2)
    for i in range(4000):
        got = oracle_case(m[i], outside[i], tab, n[i], v[i], l[i], u[i])
        want = np.zeros(16, np.float64)
        R.refh_bsdf_case(lro.ptr(m[i]), int(outside[i]), lro.ptr(n[i]), lro.ptr(v[i]), lro.ptr(l[i]), float(u[i, 0]),
                         float(u[i, 1]), float(u[i, 2]), lro.ptr(want))
        assert np.array_equal(got.astype(np.float64), want), (i, got, want)


def test_bsdf_white_furnace_and_reciprocity():
    # a white rough metal keeps all the light (multiple scattering; Kulla-Conty): the
    # directional albedo by sampling is 1 within the Monte Carlo error; eval is reciprocal
    # for reflection (test_brdf.c's checks)
    tab = lr.brdf_tables()
    rng = np.random.default_rng(3)
    nrm = np.array([0, 0, 1], np.float32)
    for rough in (0.3, 0.7, 1.0):
        m = np.array([4, 1, 1, 1, 1.5, 1.0, rough, 0.0, 0], np.float32)
        for cv in (0.3, 0.8):
            v = np.array([np.sqrt(1 - cv * cv), 0, cv], np.float32)
            acc = 0.0
            N = 20000
            for k in range(N):
                out = oracle_case(m, 1, tab, nrm, v, v, rng.uniform(0, 1, 3).astype(np.float32))
                if out[4] > 0:
                    acc += float(np.mean(out[9:12]))
            assert abs(acc / N - 1.0) < 0.02, (rough, cv, acc / N)
    for i in range(500):
        n, v, l, m, u, outside = random_cases(1, 100 + i)
        if np.dot(l[0], n[0]) <= 0.05 or np.dot(v[0], n[0]) <= 0.05 or m[0, 7] > 0:
            continue
        a = oracle_case(m[0], 1, tab, n[0], v[0], l[0], u[0])
        b = oracle_case(m[0], 1, tab, n[0], l[0], v[0], u[0])
        assert np.allclose(a[:3], b[:3], rtol=2e-3, atol=1e-6), (a[:3], b[:3])


# ---- the image -------------------------------------------------------------------------------
def test_closed_form_square_lamp_over_a_diffuse_floor():
    # A floor of albedo rho seen straight down, lit only by a square lamp of radiance Le and side
    # s at height h, centred above the pixel's point: with one bounce, the pixel reads
    # rho / pi * E, E = Le * 4 * F(s/2, s/2, h), F the corner-point integral of a rectangle
    # (a, b) at height h: (1/2)(a/A atan(b/A) + b/B atan(a/B)), A = sqrt(a^2+h^2), B = sqrt(b^2+h^2)
    rho, Le, s, h = 0.5, 10.0, 0.2, 0.5
    mats = (lr.Material * 2)()
    mats[0].type = lr.MAT_DIFFUSE
    mats[0].albedo[:] = [rho] * 3
    mats[0].ior, mats[0].roughness = 1.5, 1.0
    mats[1].type = lr.MAT_LIGHT
    mats[1].emission[:] = [Le] * 3
    floor = lr.make_quad([-5, 0, 5], [10, 0, 0], [0, 0, -10], 0)   # u x v points up
    lamp = lr.make_quad([s / 2, h, s / 2], [-s, 0, 0], [0, 0, -s], 1)   # faces down
    prims = np.stack([lamp, floor])
    cam = lr.camera_make((0, 0.3, 0), (0, 0, 0), (0, 0, 1), 0.01, 1.0)   # a tiny view of the point below
    scene = {"prims": prims, "tris": np.zeros((0, 12), np.float32), "mats": mats, "light": 0, "cam": cam, "eps": 1e-4,
             "margin": 1e-3}
    a = b = s / 2
    A, B = np.hypot(a, h), np.hypot(b, h)
    E = Le * 4 * 0.5 * (a / A * np.arctan(b / A) + b / B * np.arctan(a / B))
    lambert = rho / np.pi * E
    # the reference's diffuse material is Lambert under a rough dielectric coat (brdf.h:10-18),
    # so the exact answer is the integral of its BSDF over the lamp: midpoint quadrature,
    # 80 x 80 points, f from the oracle's bsdf_eval
    tab = lr.brdf_tables()
    up = np.array([0, 1, 0], np.float32)
    k = 80
    xs = (np.arange(k) + 0.5) / k * s - s / 2
    exact = 0.0
    for x in xs:
        for z in xs:
            d = np.array([x, h, z])
            d2 = float(d @ d)
            l = (d / np.sqrt(d2)).astype(np.float32)
            f = oracle_case(mats_row(mats[0]), 1, tab, up, up, l, np.zeros(3, np.float32))[0]
            exact += f * Le * l[1] * l[1] / d2 * (s / k) ** 2
    assert abs(exact - lambert) < 0.02 * lambert          # "not quite Lambert's" (test_scene.c:146)
    for mis in (True, False):
        img = lro.pt_image(scene, 4, 4, 0, 4096, max_depth=1, mis=mis, tab=tab, mask=lro.blue_noise())
        got = img.mean() / 4096
        assert abs(got - exact) < 0.003 * exact, (mis, got, exact)


def test_oracle_image_matches_reference_fp64():
    # the same Cornell box through the reference's own trace_image (double build) and the
    # oracle: image means within 3 sigma of their noise, and per-pixel z within bounds
    # Each batch has its own scramble (seed): consecutive chunks of one Owen-scrambled sequence
    # are not independent, and their spread would understate the error. With K batches per
    # side the z statistic is heavy-tailed (t with 2K - 2 degrees of freedom).
    need_ref(True)
    W = H = 24
    spp, K = 256, 8
    sc = lr.cornell(W, H)
    tab, mask = lr.brdf_tables(), lro.blue_noise()
    batches = [lro.pt_image(sc, W, H, 0, spp, seed=200 + k, tab=tab, mask=mask) / spp for k in range(K)]
    ours = np.mean(batches, 0)
    refs = [lro.ref_render(sc, W, H, spp, seed=100 + k) for k in range(K)]
    theirs = np.mean(refs, 0)
    var = (np.var(batches, 0, ddof=1) + np.var(refs, 0, ddof=1)) / K
    z = (ours - theirs) / np.sqrt(var + 1e-12)
    assert np.mean(np.abs(z) > 3) < 0.03 and np.abs(z).max() < 8
    sig = np.sqrt((np.var([b.mean() for b in batches], ddof=1) + np.var([r.mean() for r in refs], ddof=1)) / K)
    assert abs(ours.mean() - theirs.mean()) < 3 * sig + 1e-3 * theirs.mean()


def test_tree_walked_oracle_equals_brute_force():
    # the oracle's path tracer may walk its own LBVH (pinned against brute force) for a scene's
    # triangles: the images are the brute force's, to the bit but for rays with two hits at
    # exactly equal t (none expected on a soup)
    sc = lr.batch_scene(20000, 5)
    sc["cam"] = lr.batch_camera(9, 24, 24)
    tab, mask = lr.brdf_tables(), lro.blue_noise()
    a = lro.pt_image(dict(sc, tree=True), 24, 24, 0, 4, seed=3, tab=tab, mask=mask)
    b = lro.pt_image(dict(sc, tree=False), 24, 24, 0, 4, seed=3, tab=tab, mask=mask)
    assert a.sum() > 0
    assert np.mean(a.view(np.uint32) == b.view(np.uint32)) > 0.999


def _glow_samples(box, mis, emitters=True, n=200, seed=0):
    """the oracle's per-sample values at the floor point below a glowing shape: one sample per
    pixel of an n x n image whose view spans a few hundredths of a unit around it"""
    from glow_scenes import glow_scene
    sc, _ = glow_scene(box=box, W=n, H=n)
    sc["emitters"] = emitters
    img = lro.pt_image(sc, n, n, 0, 1, seed=seed, max_depth=8, mis=mis, tab=lr.brdf_tables(), mask=lro.blue_noise())
    return img[..., 0].astype(np.float64).ravel()


def test_glowing_ball_closed_form():
    # test_scene.c:118-153: 0.5 * 10 * (10 / 50)^2 = 0.2 with NEE and MIS, with NEE alone, and
    # with the paths alone (noisier); NEE's noise a fraction of the paths' alone
    nee_mis, nee, alone = (_glow_samples(False, m, e) for m, e in ((True, True), (False, True), (True, False)))
    for name, v, tol in (("NEE + MIS", nee_mis, 0.004), ("NEE alone", nee, 0.004), ("paths alone", alone, 0.012)):
        assert abs(v.mean() - 0.2) < tol, (name, v.mean())
    assert nee_mis.std() < 0.2 * alone.std() and nee.std() < 0.2 * alone.std(), (nee_mis.std(), nee.std(), alone.std())


def test_glowing_box_nee_agrees_with_the_paths_alone():
    # test_scene.c:155-170: a turned box; NEE's mean is the paths' alone, and its noise under half
    a, b = _glow_samples(True, True, True), _glow_samples(True, True, False)
    se = np.sqrt(a.var() / a.size) + np.sqrt(b.var() / b.size)
    assert abs(a.mean() - b.mean()) < 3 * se, (a.mean(), b.mean(), se)
    assert a.std() < 0.5 * b.std()


def test_emitter_table_rules():
    # scene_find_emitters (scene.c:1000-1013): glowing unit spheres whose axes have one length and
    # glowing boxes; a stretched ball, a lamp material, a mesh instance and a dark material are not
    from glow_scenes import _mats, axes_xform
    mats = _mats()
    fwd = np.stack([axes_xform(np.eye(3), (10, 10, 10), (0, 50, 0)), axes_xform(np.eye(3), (10, 20, 10), (0, 0, 0)),
                    axes_xform(np.eye(3), (3, 4, 5), (1, 2, 3)), axes_xform(np.eye(3), (1, 1, 1), (0, 0, 0)),
                    axes_xform(np.eye(3), (1, 1, 1), (0, 0, 0))])
    kind = np.array([lro.INST_SPHERE, lro.INST_SPHERE, lro.INST_BOX, 0, lro.INST_SPHERE])
    mat = np.array([1, 1, 1, 1, 0])
    em, slot, n = lro.find_emitters(fwd, kind, mat, mats)
    assert n == 2 and list(slot[:5]) == [0, -1, 1, -1, -1]
    e = em[:n]
    assert e[0, 0].view(np.int32) == 1 and np.allclose(e[0, 1:4], [0, 50, 0]) and abs(e[0, 13] - 10) < 1e-6
    assert abs(e[0, 14] - 4 * np.pi * 100) < 1e-3                   # 4 pi r^2
    assert abs(e[1, 14] - 8 * (4 * 5 + 3 * 5 + 3 * 4)) < 1e-3       # 8 (|a1 x a2| + |a0 x a2| + |a0 x a1|)
    assert np.allclose(e[1, 15:18], 10)
