"""GPU tests of the drop-in: the reference's own engine code drives liblr_hip.so through the
reference-side binding (little_renderer_amd/binding/gpu_frame_hip.c, built in place of
engine/app/gpu_frame.c into oracle/_ref/libref_dropin.so with oracle/dropin_harness.c).

- Images: the reference's own scenes (scene_cornell_box's boxes and spheres; instances under its
  procedural sky) through gpu_frame_open / gpu_frame_trace / gpu_frame_close, against the
  reference's own CPU trace_image on the same scene. They agree in distribution, not sample for
  sample: the device draws its plain numbers from a PCG32 stream per (pixel, sample), so that
  any split of the samples gives the same image, where trace_image runs one stream per pixel
  through its samples (the Sobol dimensions, which drive the first bounces, are the same). So
  each pixel is held to its Monte Carlo noise: K independent CPU images (seeds 1 .. K) give each
  pixel channel's mean and spread, and z = (gpu - mean) / (spread sqrt(1 + 1/K)). Glass and
  mirrors make pixels' samples heavy-tailed (caustics), so z is not Student's t there: the share
  of |z| > 3 is held to 1.5x the median share that each CPU image shows against the other K - 1
  (what two runs of the same estimator show), and the image means agree within 4 standard errors.
- With light-traced caustics (gpu_frame_trace's photons): against the reference's CPU
  trace_image with skip_caustics plus its light_trace_caustics, the same way.
- Twice on a new GPU and view: the same image to the last bit; exactly spp samples.
- A scene of instances, two instances moved (scene_move_instance + scene_update_instances): the
  binding's kept copy is brought up to date by the instances' versions and equals the CPU's image
  of the moved scene.
- Batches of scene_intersect / scene_occluded through the binding against the CPU's, ray by ray.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "libref_dropin.so")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available(), "these tests need the GPU"
    if not os.path.exists(DROPIN):
        pytest.skip("oracle/_ref/libref_dropin.so was not built (the reference was not mounted at build time)")
    lib = ctypes.CDLL(DROPIN)
    P, I = ctypes.c_void_p, ctypes.c_int
    lib.dropin_render.argtypes = [I, I, I, I, I, I, ctypes.c_double, P, P, P, P]
    lib.dropin_render.restype = I
    lib.dropin_queries.argtypes = [I, I, ctypes.c_uint64, P, P]
    lib.dropin_queries.restype = I
    return lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


K = 16


def render(L, kind, W, H, spp, move=0, photons=0.0):
    gpu, gpu2 = (np.zeros((H, W, 3), np.float32) for _ in range(2))
    cpu = np.zeros((K, H, W, 3), np.float32)
    passes = np.zeros(2, np.int32)
    rc = L.dropin_render(kind, W, H, spp, move, K, photons, ptr(gpu), ptr(gpu2), ptr(cpu), ptr(passes))
    assert rc == 0, f"the binding refused or failed (rc {rc})"
    return gpu, gpu2, cpu, passes


def z_share(img, others):
    """the share of pixel channels of img more than 3 of their standard errors from the others' mean"""
    k = others.shape[0]
    mean = others.mean(0, dtype=np.float64)
    var = others.var(0, ddof=1, dtype=np.float64) * (1 + 1 / k)
    # a pixel with no noise (the lamp seen directly, the sky): float rounding alone
    z = (img - mean) / np.sqrt(var + (1e-5 * np.abs(mean)) ** 2 + 1e-20)
    return np.mean(np.abs(z) > 3), mean, var


def compare(gpu, cpu, what):
    share, mean, var = z_share(gpu, cpu)
    controls = [z_share(cpu[k], np.delete(cpu, k, 0))[0] for k in range(K)]
    d = gpu.mean(dtype=np.float64) - mean.mean()
    se = np.sqrt(var.sum()) / var.size
    say(f"  {what}: |z|>3 in {share:.4f} of pixel channels (one CPU image against the others: median "
        f"{np.median(controls):.4f}, max {max(controls):.4f}), image means {gpu.mean():.6f} {mean.mean():.6f} "
        f"({d / se:+.2f} standard errors)")
    assert share < 1.5 * np.median(controls)
    assert abs(d) < 4 * se


@pytest.mark.parametrize("kind", [0, 1], ids=["cornell-boxes", "cornell-spheres"])
def test_gpu_frame_trace_equals_the_reference_cpu_image(L, kind):
    W = H = 64
    spp = 64
    say(f"drop-in: scene {kind} {W}x{H} spp {spp}")
    gpu, gpu2, cpu, passes = render(L, kind, W, H, spp)
    assert list(passes) == [spp, spp]
    assert np.array_equal(gpu, gpu2), "traced twice on a new GPU and view: not the same image"
    assert gpu.mean() > 0
    compare(gpu, cpu, "gpu_frame_trace vs trace_image")


def test_gpu_frame_trace_with_caustics_equals_the_reference_cpu_image(L):
    # gpu_frame_trace's photons: the device's light tracing against the reference's CPU
    # trace_image with skip_caustics plus its light_trace_caustics, on its mirror and glass balls
    W = H = 64
    spp, photons = 64, 1.0
    say(f"drop-in: scene 1 {W}x{H} spp {spp}, {photons} light paths a pixel and sample")
    gpu, gpu2, cpu, passes = render(L, 1, W, H, spp, photons=photons)
    assert list(passes) == [spp, spp]
    assert np.array_equal(gpu, gpu2), "traced twice on a new GPU and view: not the same image"
    compare(gpu, cpu, "gpu_frame_trace with caustics vs trace_image + light_trace_caustics")


def test_instanced_scene_under_the_sky_and_moved_instances(L):
    W, H, spp = 64, 48, 64
    say(f"drop-in: instances under the sky {W}x{H} spp {spp}, then two instances moved")
    gpu, gpu2, cpu, passes = render(L, 2, W, H, spp, move=1)
    assert passes[0] == spp
    assert not np.array_equal(gpu, gpu2), "the moved instances did not change the image"
    compare(gpu2, cpu, "after the move: the kept copy vs trace_image of the moved scene")


@pytest.mark.parametrize("kind", [0, 1, 2], ids=["cornell-boxes", "cornell-spheres", "instances"])
def test_batch_queries_equal_scene_intersect(L, kind):
    n = 4000
    say(f"drop-in queries: scene {kind}, {n} rays")
    out = np.zeros((n, 10), np.float32)
    occ = np.zeros((n, 2), np.float32)
    assert L.dropin_queries(kind, n, 7, ptr(out), ptr(occ)) == 0
    g, c = out[:, :5], out[:, 5:]
    both = (g[:, 0] == 1) & (c[:, 0] == 1)
    # a ray that leaves through a seam (two shapes' shared edge) may be found by one side only
    assert np.mean(g[:, 0] != c[:, 0]) < 2e-3
    same = both & (g[:, 1] == c[:, 1]) & (g[:, 2] == c[:, 2])
    trel = np.abs(g[:, 3] - c[:, 3]) / np.maximum(np.abs(c[:, 3]), 1e-6)
    # a different shape only at an equal distance (the tie rule: the first tested keeps it)
    ties = both & ~same & (trel < 1e-4)
    say(f"  found {both.mean():.3f}, same shape {same.sum()}, ties {ties.sum()}, other {(both & ~same & ~ties).sum()}")
    assert not np.any(both & ~same & ~ties)
    # one-level scenes: t within 1e-5 relative; an instance's t is found in its object space,
    # through an inverse transform the binding's library takes in double and the CPU in float
    assert np.all(trel[same] < (1e-5 if kind <= 1 else 5e-5))
    assert np.array_equal(g[same, 4], c[same, 4])          # front_face
    assert np.mean(occ[:, 0] != occ[:, 1]) < 2e-3
