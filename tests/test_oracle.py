"""CPU tests: pin the oracle (oracle/liblro.so) before trusting it.

Each check is chosen so that a plausible slip in the restatement fails it: published known
answers (PCG32), closed forms (Morton interleave), an independent construction (the binary
radix tree by recursive bit splits; PLOC as global mutual-nearest agglomeration where the
radius covers every cluster), the reference's own CPU engine compiled from its sources
(oracle/_ref, where it was built), and brute force over every triangle.
"""
import numpy as np
import pytest

from oracle import lro
import little_renderer_amd as lr


def soup(n, seed=2):
    return lr.gen_soup(n, seed)


# ---- PCG32 ------------------------------------------------------------------------------
def test_pcg32_known_answers():
    # pcg32-demo (pcg-random.org, pcg_basic.c) seeded (42, 54): its first six outputs
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert list(lro.pcg32(42, 54, 6)) == want


def test_pcg32_matches_reference_build():
    R = lro.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (no /root/reference here)")
    out = np.zeros(1000, np.uint32)
    R.refh_pcg32(0x5eedb10e, 7, 1000, lro.ptr(out))
    assert np.array_equal(out, lro.pcg32(0x5eedb10e, 7, 1000))


# ---- boxes and Morton codes -----------------------------------------------------------------
def test_tri_boxes_contain_corners_with_pad():
    t = soup(500)
    b = lro.tri_boxes(t)
    q, u, v = t[:, 0:3], t[:, 4:7], t[:, 8:11]
    corners = np.stack([q, q + u, q + v], 1)            # float32 sums, as the reference
    lo = corners.min(1) - np.float32(1e-3)
    hi = corners.max(1) + np.float32(1e-3)
    assert np.array_equal(b[:, 0:3], lo) and np.array_equal(b[:, 4:7], hi)


def interleave(x, y, z):
    """bit k of x at 3k+2, of y at 3k+1, of z at 3k: the loop the shifts-and-masks replace"""
    c = 0
    for k in range(10):
        c |= ((x >> k) & 1) << (3 * k + 2) | ((y >> k) & 1) << (3 * k + 1) | ((z >> k) & 1) << (3 * k)
    return c


def test_morton_closed_form():
    rng = np.random.default_rng(1)
    n = 300
    q = rng.integers(0, 1023, size=(n, 3))
    # boxes whose centres land on known quantization cells: centre = (q + 0.5) / 1024 in [0,1)^3
    c = ((q + 0.5) / 1024.0).astype(np.float32)
    boxes = np.zeros((n, 8), np.float32)
    boxes[:, 0:3] = c
    boxes[:, 4:7] = c
    lo = np.zeros(3, np.float32)
    ext = np.ones(3, np.float32)
    codes = lro.morton(boxes, lo, ext)
    want = [interleave(*map(int, row)) for row in q]
    assert list(codes) == want
    # clamping: below the low corner -> 0, at or past the far side -> 1022 (0.999 * 1024)
    edge = np.zeros((2, 8), np.float32)
    edge[0, 0:3] = edge[0, 4:7] = -5.0
    edge[1, 0:3] = edge[1, 4:7] = 7.0
    e = lro.morton(edge, lo, ext)
    assert e[0] == 0 and e[1] == interleave(1022, 1022, 1022)


def test_inv_extent_double_then_float():
    lo = np.array([0.25, -3.0, 1.0], np.float32)
    hi = np.array([0.75, 4.0, 1.0], np.float32)         # flat z: 0
    ext = lro.inv_extent(lo, hi)
    assert ext[0] == np.float32(2.0) and ext[1] == np.float32(1.0 / 7.0) and ext[2] == 0.0


def test_sort_is_lexicographic_code_then_item():
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 64, size=5000).astype(np.uint32)   # many equal codes
    sc, si = lro.sort(codes)
    order = np.lexsort((np.arange(5000), codes))
    assert np.array_equal(si, order.astype(np.uint32)) and np.array_equal(sc, codes[order])


# ---- LBVH ------------------------------------------------------------------------------------
def radix_tree(sc):
    """The binary radix tree over the augmented keys (code << 32 | position), built by splitting
    each range at the highest bit where its first and last keys differ, numbered as Karras
    (root 0; a left child [l, g] is internal g, a right child [g+1, r] internal g+1; leaf j is
    n-1+j). Returns the children arrays (a, b) of the n-1 internal nodes."""
    n = len(sc)
    keys = [(int(c) << 32) | j for j, c in enumerate(sc)]
    A = [None] * max(n - 1, 0)
    B = [None] * max(n - 1, 0)

    def node(l, r, idx):
        x = keys[l] ^ keys[r]
        bit = x.bit_length() - 1
        g = l
        while not (keys[g + 1] >> bit) & 1:
            g += 1
        lc = n - 1 + g if l == g else g
        rc = n - 1 + g + 1 if g + 1 == r else g + 1
        A[idx], B[idx] = lc, rc
        if l != g:
            node(l, g, g)
        if g + 1 != r:
            node(g + 1, r, g + 1)

    if n > 1:
        node(0, n - 1, 0)
    return A, B


@pytest.mark.parametrize("n,codes_range", [(1, 1 << 30), (2, 1 << 30), (3, 1 << 30), (57, 1 << 30), (200, 16),
                                           (513, 1 << 30), (1000, 3)])
def test_lbvh_is_the_binary_radix_tree(n, codes_range):
    rng = np.random.default_rng(n)
    codes = rng.integers(0, codes_range, size=n).astype(np.uint32)
    boxes = lro.tri_boxes(soup(n, seed=n))
    sc, si = lro.sort(codes)
    nodes, aux, depth = lro.lbvh(sc, si, boxes)
    A, B = radix_tree(sc)
    assert list(nodes[: n - 1, 3]) == A and list(nodes[: n - 1, 7]) == B
    # leaves: box of the sorted item, a = its slot, b = one shape
    leaves = nodes[n - 1:]
    assert np.array_equal(leaves[:, 7], np.full(n, 0x80000001, np.uint32))
    assert np.array_equal(leaves[:, 3], np.arange(n, dtype=np.uint32))
    assert np.array_equal(aux, si)
    assert np.array_equal(leaves.view(np.float32)[:, [0, 1, 2, 4, 5, 6]], boxes[si][:, [0, 1, 2, 4, 5, 6]])
    ok, cost, d = lro.check_tree(nodes, n)
    assert ok and d == depth


def test_lbvh_internal_boxes_are_unions():
    n = 777
    boxes = lro.tri_boxes(soup(n, seed=9))
    lo, hi = lro.union(boxes)
    codes = lro.morton(boxes, lo, lro.inv_extent(lo, hi))
    sc, si = lro.sort(codes)
    nodes, aux, depth = lro.lbvh(sc, si, boxes)
    f = nodes.view(np.float32)

    def box(i):
        a, b = nodes[i, 3], nodes[i, 7]
        if b & 0x80000000:
            return f[i, 0:3], f[i, 4:7]
        l0, h0 = box(a)
        l1, h1 = box(b)
        return np.minimum(l0, l1), np.maximum(h0, h1)

    l, h = box(0)
    assert np.array_equal(l, f[0, 0:3]) and np.array_equal(h, f[0, 4:7])
    assert np.array_equal(f[0, 0:3], lo) and np.array_equal(f[0, 4:7], hi)


@pytest.mark.parametrize("n,codes_range", [(2, 1 << 30), (5, 4), (333, 1 << 30), (4096, 1 << 30), (5000, 7),
                                           (20000, 1 << 12)])
def test_bottom_up_equals_karras(n, codes_range):
    rng = np.random.default_rng(n + 11)
    codes = rng.integers(0, codes_range, size=n).astype(np.uint32)
    boxes = lro.tri_boxes(soup(n, seed=n + 1))
    sc, si = lro.sort(codes)
    a = lro.lbvh(sc, si, boxes, node0=5, ref0=9, ref_add=3)
    b = lro.lbvh(sc, si, boxes, node0=5, ref0=9, ref_add=3, bottom_up=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


# ---- PLOC ------------------------------------------------------------------------------------
def half_area32(lo, hi):
    e = (hi - lo).astype(np.float32)
    a = np.float32(e[0] * e[1])
    b = np.float32(e[1] * e[2])
    c = np.float32(e[2] * e[0])
    return np.float32(np.float32(a + b) + c)


def ploc_global(boxes_sorted, n, live=None):
    """Agglomerative clustering where every cluster sees every other (n <= R + 1 = 17, so the
    radius of 16 covers the whole list): each round, each cluster's nearest over the whole list
    (ties: the lower pair of positions), mutual pairs merge into the lower one's place, the
    list keeps its order; a merge makes internal node (n-2) - (merges before it). Returns the
    node array as trace_ploc.wgsl lays it out; `live` (a list) gets each round's cluster count."""
    bx = {}
    for j in range(n):
        bx[n - 1 + j] = (boxes_sorted[j, 0:3].copy(), boxes_sorted[j, 4:7].copy())
    nodes = np.zeros((2 * n - 1, 8), np.uint32)
    f = nodes.view(np.float32)
    lst = [n - 1 + j for j in range(n)]
    merges = 0

    def final(slot, x):
        lo, hi = bx[x]
        f[slot, 0:3] = lo
        f[slot, 4:7] = hi
        if x < n - 1:
            nodes[slot, 3], nodes[slot, 7] = 1 + 2 * x, 2 + 2 * x
        else:
            nodes[slot, 3], nodes[slot, 7] = x - (n - 1), 0x80000001

    while len(lst) > 1:
        nc = len(lst)
        if live is not None:
            live.append(nc)
        nn = []
        for i in range(nc):
            best = None
            for j in range(nc):
                if j == i:
                    continue
                lo = np.minimum(bx[lst[i]][0], bx[lst[j]][0])
                hi = np.maximum(bx[lst[i]][1], bx[lst[j]][1])
                key = (half_area32(lo, hi), min(i, j), max(i, j))
                if best is None or key < best[0]:
                    best = (key, j)
            nn.append(best[1])
        out = []
        for i in range(nc):
            j = nn[i]
            mu = nn[j] == i
            if mu and i > j:
                continue
            if mu:
                k = (n - 2) - merges
                merges += 1
                c, d = lst[i], lst[j]
                bx[k] = (np.minimum(bx[c][0], bx[d][0]), np.maximum(bx[c][1], bx[d][1]))
                final(1 + 2 * k, c)
                final(2 + 2 * k, d)
                if k == 0:
                    f[0, 0:3], f[0, 4:7] = bx[0]
                    nodes[0, 3], nodes[0, 7] = 1, 2
                out.append(k)
            else:
                out.append(lst[i])
        lst = out
    return nodes


@pytest.mark.parametrize("n", [3, 4, 9, 17])
def test_ploc_small_is_global_agglomeration(n):
    for seed in range(5):
        t = soup(n, seed=100 + seed)
        nodes, aux, sc, si, depth = lro.build_tris(t, ploc_=True)
        boxes = lro.tri_boxes(t)
        counts = []
        want = ploc_global(boxes[si], n, counts)
        assert np.array_equal(nodes, want), (n, seed)
        assert np.array_equal(aux, si)
        # the clusters summed over the rounds (the build's traffic count) agree too
        assert lro.ploc(sc, si, boxes, live=True)[4] == sum(counts), (n, seed)


def test_ploc_ties_go_to_the_lower_pair():
    # 6 identical boxes: every distance ties, so each cluster's nearest is its lower neighbour
    # pair ((0,1) for 0 and 1, (1,2)... -> the lowest positions): 0-1 merge first
    n = 6
    boxes = np.tile(np.array([0, 0, 0, 0, 1, 1, 1, 0], np.float32), (n, 1))
    sc = np.zeros(n, np.uint32)
    si = np.arange(n, dtype=np.uint32)
    nodes, aux, depth, rounds, live = lro.ploc(sc, si, boxes, live=True)
    want = ploc_global(boxes, n)
    assert np.array_equal(nodes, want)
    # one pair a round (the lowest): n, n - 1, ..., 2 clusters live
    assert rounds == n - 1 and live == n * (n + 1) // 2 - 1
    ok, cost, d = lro.check_tree(nodes, n)
    assert ok and d == depth


@pytest.mark.parametrize("n", [3, 100, 2000, 30000])
def test_ploc_is_one_tree(n):
    t = soup(n, seed=7)
    nodes, aux, sc, si, depth = lro.build_tris(t, ploc_=True)
    ok, cost, d = lro.check_tree(nodes, n)
    assert ok and d == depth
    assert np.array_equal(np.sort(aux), np.arange(n, dtype=np.uint32))


def sphere_mesh(nu=200, nv=100):
    tris = []
    for i in range(nu):
        for j in range(nv):
            th, ph = np.pi * j / nv, 2 * np.pi * i / nu

            def p(th, ph):
                return np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])
            a, b, c = p(th, ph), p(th + np.pi / nv, ph), p(th, ph + 2 * np.pi / nu)
            tris.append(np.concatenate([a, [0], b - a, [0], c - a, [0]]))
    return np.array(tris, np.float32)


def test_ploc_cheaper_than_lbvh_on_a_mesh():
    # on a mesh's surface PLOC's trees cost less than the LBVH's over the same keys (the
    # reference measured 8-9% on its models, docs/STATUS.md:706-710); on a uniform soup of tiny
    # boxes the Morton splits are already near the best and PLOC is not cheaper
    t = sphere_mesh()
    n = t.shape[0]
    pn = lro.build_tris(t, ploc_=True)[0]
    ln = lro.build_tris(t, ploc_=False)[0]
    assert lro.check_tree(pn, n)[1] < 0.98 * lro.check_tree(ln, n)[1]


# ---- the CPU's SAH tree against the reference's own build ----------------------------------------
@pytest.mark.parametrize("n", [1, 5, 100, 3000])
def test_sah_tree_matches_reference(n):
    R = lro.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (no /root/reference here)")
    t = soup(n, seed=21)
    boxes = lro.tri_boxes(t)
    mine = lro.Sah(boxes, 4, 16, wide=False)
    nodes, aux, depth = mine.put_nodes()
    rn = np.zeros((2 * n - 1, 8), np.uint32)
    rr = np.zeros(n, np.int32)
    rd = np.zeros(1, np.int32)
    rc = np.zeros(1, np.float64)
    nn = R.refh_bvh_tris(lro.ptr(t), n, 4, 16, lro.ptr(rn), lro.ptr(rr), lro.ptr(rd), lro.ptr(rc))
    assert nn == nodes.shape[0]
    assert np.array_equal(rn[:nn], nodes) and np.array_equal(rr.astype(np.uint32), aux)
    assert rd[0] == depth
    assert abs(rc[0] - mine.cost()) <= 1e-6 * rc[0]


# ---- traversal -------------------------------------------------------------------------------
def random_rays(n, seed, tmin=0.0, tris=None):
    """rays from random points around the unit cube; with tris, half of them aimed at a random
    point inside a random triangle (so that many hit)"""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.2, 1.2, size=(n, 3)).astype(np.float32)
    tgt = rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32)
    if tris is not None:
        k = rng.integers(0, tris.shape[0], size=n // 2)
        a, b = rng.uniform(0, 0.5, size=(2, n // 2, 1))
        tgt[: n // 2] = tris[k, 0:3] + a * tris[k, 4:7] + b * tris[k, 8:11]
    d = (tgt - o).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = 1e30
    return r


def compare_ids(h, ref, second, tol=1e-5):
    """identical hit ids except on ties (the two nearest candidates within tol relative);
    t to the bit where the ids agree (the same triangle test). Returns the tie count."""
    ih, ir = lro.hit_ids(h), lro.hit_ids(ref)
    tie = (ir != 0xFFFFFFFF) & (np.abs(second - ref[:, 0]) <= tol * np.abs(ref[:, 0]))
    bad = (ih != ir) & ~tie
    assert not bad.any(), f"{bad.sum()} rays differ: first {np.nonzero(bad)[0][:5]}"
    same = (ih == ir) & (ir != 0xFFFFFFFF)
    assert np.array_equal(h[same], ref[same])
    return int(tie.sum())


@pytest.mark.parametrize("ploc", [False, True])
def test_gpu_layout_walk_matches_brute_force(ploc):
    n = 3000
    t = soup(n, seed=5)
    nodes, aux, sc, si, depth = lro.build_tris(t, ploc_=ploc)
    rays = random_rays(4000, 6, tris=t)
    h = lro.trace_gpu_layout(nodes, 0, aux, t, rays)
    b, second = lro.trace_brute(t, rays)
    assert (lro.hit_ids(b) != 0xFFFFFFFF).sum() > 500
    compare_ids(h, b, second)
    # any-hit: a hit exactly where the closest-hit walk has one
    a = lro.trace_gpu_layout(nodes, 0, aux, t, rays, any_hit=True)
    assert np.array_equal(lro.hit_ids(a) != 0xFFFFFFFF, lro.hit_ids(h) != 0xFFFFFFFF)


def test_sah_walks_match_brute_force_and_reference():
    n = 3000
    t = soup(n, seed=8)
    boxes = lro.tri_boxes(t)
    rays = random_rays(4000, 9, tris=t)
    b, second = lro.trace_brute(t, rays)
    for wide in (False, True):
        h = lro.Sah(boxes, 4, 16, wide=wide).trace(t, rays)
        compare_ids(h, b, second)
        R = lro.ref()
        if R is not None:
            rh = np.zeros((rays.shape[0], 4), np.float32)
            R.refh_trace_tris(lro.ptr(t), n, 1 if wide else 0, lro.ptr(rays), rays.shape[0], 0, lro.ptr(rh))
            assert np.array_equal(rh.view(np.uint32), h.view(np.uint32))
