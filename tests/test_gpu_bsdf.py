"""GPU test of the engine's BSDF at the kernel level: lrh_bsdf_cases (the same bsdf_make / eval /
pick / sample_lobe code k_shade runs) against the oracle's C BSDF on 65,536 random cases, as the
reference checks its WGSL BSDF against its C one (tests/test_gpu_brdf.c:21-90): six kinds of
material (rough dielectric, metal, glass, a mix, smooth plastic, smooth glass or metal), viewers
cosine-distributed about the normal, light directions uniform on the sphere. Evaluation (f, pdf)
and sampling (the decision, delta and through flags, the direction, weight and pdf) must agree
to 2e-3 relative (directions to 3e-3 in L1), with at most N / 2000 cases differing in all: the
device's sqrt / pow / trig may round differently from glibc's in the last bits.
"""
import numpy as np
import pytest

import little_renderer_amd as lr
from oracle import lro

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N = 65536


def cases(seed=11):
    rng = np.random.default_rng(seed)
    c = np.zeros((N, 17), np.float32)
    kind = np.arange(N) % 6
    u = lambda: rng.random(N)
    c[:, 0], c[:, 1], c[:, 2] = u(), u(), u()
    mix = u()
    c[:, 3] = np.where(kind == 1, 1.0, np.where(kind == 3, mix, 0.0))
    smooth = np.where(u() < 0.5, 0.0, 0.01)
    c[:, 4] = np.where(kind >= 4, smooth, 0.03 + 0.97 * u())
    c[:, 5] = 1.2 + 0.8 * u()
    mix2 = u()
    c[:, 6] = np.where(kind == 2, 1.0, np.where(kind == 3, mix2, np.where(kind == 5, 1.0, 0.0)))
    metal = (kind == 5) & (np.arange(N) % 12 == 5)
    c[metal, 3], c[metal, 6] = 1.0, 0.0
    c[:, 7] = (u() < 0.7).astype(np.float32)
    # v: cosine-weighted about +z; l: uniform in the cube, normalised
    r, phi = np.sqrt(u()), 2 * np.pi * u()
    v = np.stack([r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1 - r * r))], 1)
    l = rng.uniform(-1, 1, (N, 3))
    l /= np.linalg.norm(l, axis=1, keepdims=True)
    c[:, 8:11], c[:, 11:14] = v, l
    c[:, 14], c[:, 15], c[:, 16] = u(), u(), u()
    return c


def rel(a, b):
    return np.abs(a - b) / np.maximum(1e-3, np.maximum(np.abs(a), np.abs(b)))


def test_device_bsdf_matches_oracle_on_65536_cases():
    assert torch.cuda.is_available(), "this test needs the GPU"
    ctx = lr.Context(0)
    c = cases()
    want = lro.bsdf_cases(c, lr.brdf_tables())
    dc = torch.from_numpy(c).cuda()
    out = torch.zeros((N, 16), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.bsdf_cases(dc, out)
    got = out.cpu().numpy()
    ctx.close()
    assert np.isfinite(got).all()
    bad_eval = rel(want[:, 0:3], got[:, 0:3]).max(1) > 2e-3
    bad_pdf = rel(want[:, 3], got[:, 3]) > 2e-3
    ok_w, ok_g = want[:, 4] > 0.5, got[:, 4] > 0.5
    flags = (ok_w != ok_g) | (ok_w & ((want[:, 13] != got[:, 13]) | (want[:, 14] != got[:, 14])))
    drawn = ok_w & ~flags
    bad_dir = drawn & (np.abs(want[:, 6:9] - got[:, 6:9]).sum(1) > 3e-3)
    bad_w = drawn & ~bad_dir & ((rel(want[:, 9:12], got[:, 9:12]).max(1) > 2e-3) | (rel(want[:, 12], got[:, 12]) > 2e-3))
    bad = int(bad_eval.sum() + bad_pdf.sum() + flags.sum() + bad_dir.sum() + bad_w.sum())
    print(f"bsdf: eval {bad_eval.sum()}, pdf {bad_pdf.sum()}, drawn {drawn.sum()}: flags {flags.sum()}, "
          f"directions {bad_dir.sum()}, weights {bad_w.sum()}; lobes agree {np.mean(want[:, 5] == got[:, 5]):.5f}",
          flush=True)
    # every kind of case is exercised: deltas, transmissions and rough samples all occur
    assert (want[drawn, 13] > 0.5).sum() > 1000 and (want[drawn, 14] > 0.5).sum() > 1000
    assert drawn.sum() > N // 2
    assert bad <= N // 2000
