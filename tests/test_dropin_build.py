"""CPU tests of the reference-side binding (little_renderer_amd/binding/gpu_frame_hip.c): it
compiles cleanly against the reference's own headers (its definitions checked against the
prototypes of engine/app/gpu_frame.h), and the engine built with it (oracle/_ref/libref_dropin.so:
the reference's CPU engine, the binding in place of engine/app/gpu_frame.c, linked against
liblr_hip.so) loads and exports what callers use. Skipped where the reference is not mounted."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/engine"
BINDING = os.path.join(ROOT, "little_renderer_amd", "binding")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "libref_dropin.so")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's sources are not mounted")


def test_binding_compiles_against_the_reference_headers():
    for defs in ([], ["-DLR_DOUBLE"]):       # the float build and the double one
        r = subprocess.run(["gcc", "-std=gnu11", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                            *defs, f"-I{REF}", f"-I{os.path.join(ROOT, 'include')}", f"-I{BINDING}",
                            os.path.join(BINDING, "gpu_frame_hip.c")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_engine_with_the_binding_links_and_exports_the_frame_api():
    if not os.path.exists(DROPIN):
        r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "dropin"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
    L = ctypes.CDLL(DROPIN)
    for name in ("gpu_frame_open", "gpu_frame_close", "gpu_frame_trace", "gpu_frame_raster", "gpu_frame_passes",
                 "gpu_scene_intersect", "gpu_scene_occluded", "scene_intersect", "trace_image", "dropin_render"):
        assert hasattr(L, name), name
    # the binding calls the library, not a copy of it: liblr_hip.so is a needed library
    r = subprocess.run(["readelf", "-d", DROPIN], capture_output=True, text=True)
    assert "liblr_hip.so" in r.stdout
