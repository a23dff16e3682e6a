"""CPU tests of the oracle's denoiser (oracle/lro_denoise.c) and guide (lro_guide), pinned by what
the mathematics fixes rather than by retyping their formulas:

- the half packing equals IEEE-754 binary16 round-to-nearest-even (numpy's float16) on random
  floats, subnormals, ties and the edges of the range;
- the octahedral normal map round-trips every direction to within its 16-bit grid's step;
- a constant image (one flat surface) passes every pass unchanged;
- one pass over an impulse on a smooth, noisy surface spreads it by the B3 spline's weights
  (1, 4, 6, 4, 1) / 16 in each direction;
- edges stop it: no light crosses a normal edge, a distance edge, or into a lamp's pixels, and a
  lamp's pixels pass as they are;
- the guide of a camera looking down at a diffuse floor: the floor's normal, the distance along
  each pixel's centre ray (height over the ray's cosine), the floor's albedo; through a mirror,
  the surface the mirror shows, the albedo the product of the two.
"""
import ctypes

import numpy as np

import little_renderer_amd as lr
from oracle import lro


def test_half_packing_is_ieee_round_to_nearest_even():
    rng = np.random.default_rng(1)
    xs = np.concatenate([(rng.standard_normal(20000) * 10.0 ** rng.integers(-9, 5, 20000)).astype(np.float32),
                         np.array([0.0, -0.0, 1.0, 65504.0, 65519.0, 65520.0, 6e4, 2 ** -14, 2 ** -24, 2 ** -25,
                                   3 * 2 ** -26, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 2048 + 1, 2048 + 3], np.float32)])
    L = lro.lib()
    got = np.array([L.lro_f2h(float(x)) for x in xs], np.uint16)
    want = xs.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    back = np.array([L.lro_h2f(int(h)) for h in want], np.float32)
    assert np.array_equal(back, want.view(np.float16).astype(np.float32))


def test_octahedral_normals_round_trip():
    rng = np.random.default_rng(2)
    n = rng.standard_normal((5000, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n = np.concatenate([n, np.eye(3, dtype=np.float32), -np.eye(3, dtype=np.float32)])
    L = lro.lib()
    p = np.zeros(2, np.uint16)
    o = np.zeros(3, np.float32)
    worst = 0.0
    for v in n:
        v = np.ascontiguousarray(v)
        L.lro_oct_pack(lro.ptr(v), lro.ptr(p))
        L.lro_oct_unpack(lro.ptr(p), lro.ptr(o))
        worst = max(worst, float(np.linalg.norm(o.astype(np.float64) - v.astype(np.float64))))   # ~ the angle
    assert worst < 1e-4      # a 2^-15 grid on the octahedron: a few 1e-5 radians


def flat(W, H, rgb=(0.5, 0.5, 0.5), albedo=(0.5, 0.5, 0.5), z=2.0, n=(0, 0, 1), kind=0.0):
    g = np.zeros((H, W, 8), np.float32)
    g[..., 0:3], g[..., 3], g[..., 4:7], g[..., 7] = n, z, albedo, kind
    a = np.zeros((H, W, 4), np.float32)
    a[..., 0:3], a[..., 3] = np.asarray(rgb, np.float32) * 16, 16
    m = np.zeros((H, W, 2), np.float32)
    return a, m, g


def test_a_constant_image_passes_unchanged():
    W, H = 40, 30
    a, m, g = flat(W, H, rgb=(0.3, 0.4, 0.5), albedo=(0.6, 0.8, 1.0))
    lum = 0.2126 * 0.3 / 0.6 + 0.7152 * 0.4 / 0.8 + 0.0722 * 0.5 / 1.0
    m[..., 0], m[..., 1] = 16 * lum, 16 * lum * lum           # no spread: variance 0
    out = lro.denoise(a, m, g, passes=5)
    want = np.float32([0.3, 0.4, 0.5]).astype(np.float16).astype(np.float32)
    # the irradiance packed as halves, times the albedo again
    irr = (np.float32([0.3, 0.4, 0.5]) / np.float32([0.6, 0.8, 1.0])).astype(np.float16).astype(np.float32)
    assert np.allclose(out[..., 0:3], irr * np.float32([0.6, 0.8, 1.0]), rtol=0, atol=1e-6)
    assert np.all(out[..., 3] == 1.0)
    assert np.allclose(out[..., 0:3], want, rtol=1e-3)


def impulse(W, H, var=1e6):
    """a dark surface with one bright pixel in the middle, every pixel's variance `var`"""
    a, m, g = flat(W, H, rgb=(0, 0, 0), albedo=(1, 1, 1))
    cy, cx = H // 2, W // 2
    a[cy, cx, 0:3] = 16 * 1.0
    # moments giving the variance of the mean var: m2 / n - (m1 / n)^2 = var * n
    n = 16.0
    l = np.zeros((H, W), np.float32)
    l[cy, cx] = 1.0
    m[..., 0] = n * l
    m[..., 1] = n * (var * n + l * l)
    return a, m, g, cy, cx


def test_one_pass_spreads_an_impulse_by_the_b3_spline():
    W = H = 21
    a, m, g, cy, cx = impulse(W, H)
    out = lro.denoise(a, m, g, passes=1)[..., 0]
    # with a variance far above the differences the luminance weights are 1 to ~1e-4, so pixel
    # (cy + dy, cx + dx) reads (the impulse's weight at its place) / (that pixel's weight sum, 1)
    b3 = np.array([1, 4, 6, 4, 1], np.float64) / 16
    k = np.outer(b3, b3)
    assert np.allclose(out[cy - 2:cy + 3, cx - 2:cx + 3], k, rtol=2e-3, atol=1e-6)
    mask = np.ones_like(out, bool)
    mask[cy - 2:cy + 3, cx - 2:cx + 3] = False
    assert np.all(out[mask] == 0.0)


def test_edges_stop_the_filter_and_lamps_pass():
    W = H = 24
    a, m, g, cy, cx = impulse(W, H)
    ref = lro.denoise(a, m, g, passes=3)[..., 0]
    assert ref[cy, cx + 3] > 0        # without edges the light reaches three pixels over
    # the right half turned 90 degrees: no light crosses into it
    g2 = g.copy()
    g2[:, cx + 1:, 0:3] = (1, 0, 0)
    out = lro.denoise(a, m, g2, passes=3)[..., 0]
    assert np.all(out[:, cx + 1:] == 0.0) and out[cy, cx - 1] > 0
    # the right half far behind (a distance edge)
    g3 = g.copy()
    g3[:, cx + 1:, 3] = 50.0
    out = lro.denoise(a, m, g3, passes=3)[..., 0]
    assert np.all(out[:, cx + 1:] < 1e-20)
    # the right half a lamp: passes as it is (its own value), takes nothing
    g4 = g.copy()
    g4[:, cx + 1:, 7] = 1.0
    a4 = a.copy()
    a4[:, cx + 1:, 0:3] = 16 * 7.0
    out4 = lro.denoise(a4, m, g4, passes=3)
    assert np.allclose(out4[:, cx + 1:, 0:3], 7.0) and np.all(out4[:, :cx + 1, 0] < 1.0 + 1e-6)


def test_guide_closed_forms():
    mats = (lr.Material * 3)()
    mats[0].type = lr.MAT_DIFFUSE
    mats[0].albedo[:] = [0.4, 0.5, 0.6]
    mats[0].ior, mats[0].roughness = 1.5, 1.0
    mats[1].type = lr.MAT_MIRROR
    mats[1].albedo[:] = [0.9, 0.8, 0.7]
    mats[1].ior, mats[1].metallic = 1.5, 1.0
    mats[2].type = lr.MAT_LIGHT
    mats[2].emission[:] = [5, 5, 5]
    floor = lr.make_quad([-50, 0, 50], [100, 0, 0], [0, 0, -100], 0)      # up
    h = 3.0
    W = H = 16
    cam = lr.camera_make((0, h, 0), (0, 0, 0), (0, 0, 1), 30.0, 1.0)
    sc = {"prims": np.stack([floor]), "tris": np.zeros((0, 12), np.float32), "mats": mats, "light": -1, "cam": cam,
          "eps": 1e-4, "margin": 1e-3}
    g = lro.guide(sc, W, H, lr.brdf_tables())
    assert np.all(g[..., 7] == 0) and np.allclose(g[..., 0:3], [0, 1, 0]) and np.allclose(g[..., 4:7], [0.4, 0.5, 0.6])
    # the distance: height over the cosine of each centre ray with the vertical
    c = np.asarray(cam, np.float64).reshape(4, 3)
    ys, xs = np.mgrid[0:H, 0:W]
    d = c[1] + c[2] * ((xs[..., None] + 0.5) / W) - c[3] * ((ys[..., None] + 0.5) / H) - c[0]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    assert np.allclose(g[..., 3], h / -d[..., 1], rtol=1e-5)
    # a mirror above the floor, facing it, and a camera between looking up at the mirror: the guide
    # is the floor seen in it (normal up, the two legs' distance, the albedos' product)
    mirror = lr.make_quad([-50, 10, -50], [100, 0, 0], [0, 0, 100], 1)    # faces down
    cam2 = lr.camera_make((0, h, 0), (0, 10, 0), (0, 0, 1), 20.0, 1.0)
    sc2 = dict(sc, prims=np.stack([floor, mirror]), cam=cam2)
    g2 = lro.guide(sc2, W, H, lr.brdf_tables())
    assert np.all(g2[..., 7] == 0) and np.allclose(g2[..., 0:3], [0, 1, 0], atol=1e-6)
    assert np.allclose(g2[..., 4:7], np.float32([0.4, 0.5, 0.6]) * np.float32([0.9, 0.8, 0.7]))
    d2 = c2 = np.asarray(cam2, np.float64).reshape(4, 3)
    d2 = c2[1] + c2[2] * ((xs[..., None] + 0.5) / W) - c2[3] * ((ys[..., None] + 0.5) / H) - c2[0]
    d2 /= np.linalg.norm(d2, axis=-1, keepdims=True)
    assert np.allclose(g2[..., 3], (10 - h + 10) / d2[..., 1], rtol=1e-5)
    # the lamp seen directly: kind 1, albedo 1
    lamp = lr.make_quad([-1, 10, -1], [2, 0, 0], [0, 0, 2], 2)
    sc3 = dict(sc, prims=np.stack([floor, lamp]), cam=cam2)
    g3 = lro.guide(sc3, W, H, lr.brdf_tables())
    mid = g3[H // 2, W // 2]
    assert mid[7] == 1 and np.allclose(mid[4:7], 1) and abs(mid[3] - (10 - h) / d2[H // 2, W // 2, 1]) < 1e-4
