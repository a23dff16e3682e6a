"""GPU test of lrh_build_noalloc, the build whose scratch is the library's own fixed device array
(no allocation): called through the C ABI, its trees equal lrh_build's bit for bit, for both
builders, on two calls in a row, the second with fewer items (the array keeps the first call's
contents); items above its maximum are refused with an error."""
import ctypes
import sys

import numpy as np
import pytest

import little_renderer_amd as lr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MAX_ITEMS = 65536     # LRH_NOALLOC_MAX_ITEMS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build_both(ctx, n, seed, builder):
    L = lr.lib()
    tris = torch.from_numpy(lr.gen_soup(n, seed)).cuda()
    boxes = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.tri_boxes(tris, boxes)
    out = []
    for fn in (L.lrh_build, L.lrh_build_noalloc):
        nodes = torch.zeros((2 * n - 1, 8), dtype=torch.int32, device="cuda")
        aux = torch.zeros(n, dtype=torch.int32, device="cuda")
        srt = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        st = lr.BuildStats()
        rc = fn(ctx._c, lr._ptr(boxes), n, None, builder, 0, 0, 0, lr._ptr(nodes), lr._ptr(aux), lr._ptr(srt),
                ctypes.byref(st))
        assert rc == 0, L.lrh_last_error().decode()
        ctx.sync()
        out.append((nodes.cpu().numpy(), aux.cpu().numpy(), srt.cpu().numpy(), st.depth, st.rounds, st.clusters))
    return out


@pytest.mark.parametrize("builder", [lr.LBVH, lr.PLOC], ids=["lbvh", "ploc"])
def test_noalloc_build_equals_build(builder):
    assert torch.cuda.is_available(), "this test needs the GPU"
    ctx = lr.Context(0)
    try:
        for n, seed in ((MAX_ITEMS, 5), (3001, 6)):        # two calls in a row, the second smaller
            say(f"noalloc build: builder {builder} n {n}")
            (a_nodes, a_aux, a_srt, a_d, a_r, a_c), (b_nodes, b_aux, b_srt, b_d, b_r, b_c) = build_both(ctx, n, seed,
                                                                                                       builder)
            assert np.array_equal(a_nodes, b_nodes)
            assert np.array_equal(a_aux, b_aux)
            assert np.array_equal(a_srt, b_srt)
            assert (a_d, a_r, a_c) == (b_d, b_r, b_c)
        # above the maximum: refused, nothing launched
        L = lr.lib()
        rc = L.lrh_build_noalloc(ctx._c, ctypes.c_void_p(16), MAX_ITEMS + 1, None, builder, 0, 0, 0,
                                 ctypes.c_void_p(16), ctypes.c_void_p(16), None, None)
        assert rc == -4 and b"at most" in L.lrh_last_error()
    finally:
        ctx.close()
