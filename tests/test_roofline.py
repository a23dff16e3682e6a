"""The roofline byte formulas: each against a count made by hand for a tiny case, so that a
dropped or doubled term shows."""
from little_renderer_amd import roofline as rf


def test_walk_bytes_by_hand():
    # 3 rays, 10 wide nodes, 4 triangles: 3 x (32 + 16) + 10 x 64 + 4 x (48 + 4)
    assert rf.walk_bytes(3, 10, 4, rf.NODE_BYTES_WIDE) == 144 + 640 + 208
    # the reference's binary node is 32 bytes
    assert rf.walk_bytes(1, 1, 0, rf.NODE_BYTES_BINARY) == 48 + 32
    assert rf.walk_bytes(0, 0, 0, 64) == 0


def test_lbvh_bytes_by_hand():
    # bounds 32, morton 32 + 8, sort 4 x (8 + 8), lbvh 4 + 32 + 32 + 4 + (32 + 32)
    per = 32 + 40 + 64 + 136
    assert rf.lbvh_bytes(1) == per
    assert rf.lbvh_bytes(1000) == 1000 * per
    assert rf.lbvh_bytes(10, with_bounds=False) == 10 * (per - 32)


def test_ploc_bytes_by_hand():
    # 3 primitives, 5 live clusters over the rounds: bounds 32 + morton 40 + sort 64 and init
    # 80 per primitive, 40 per live cluster, 108 for each of the 2 merges
    assert rf.ploc_bytes(3, 5) == 3 * (32 + 40 + 64 + 80) + 5 * 40 + 2 * 108
    assert rf.ploc_bytes(3, 5, with_bounds=False) == rf.ploc_bytes(3, 5) - 3 * 32
    assert rf.ploc_bytes(1, 0) == 32 + 40 + 64 + 80


def test_roofline_fraction():
    gbs, frac = rf.roofline(8e9, 1.0)           # 8 GB in a millisecond: 8000 GB/s, the peak
    assert abs(gbs - 8000.0) < 1e-6 and abs(frac - 1.0) < 1e-12
    assert rf.roofline(1, 0.0) == (0.0, 0.0)


def test_shade_bytes_by_hand():
    # an entry's state 4 + 8 + 32 + 16 + 16 + 16 + 8 = 100 read, its triangle 48 and material 48;
    # a path that goes on writes 32 + 4 + 8 + 16 + 16 + 8 = 84, a shadow ray 32 + 4 + 16 + 16 = 68
    assert rf.SHADE_IN == 4 + 8 + 32 + 16 + 16 + 16 + 8
    assert rf.SHADE_NEXT == 32 + 4 + 8 + 16 + 16 + 8 and rf.SHADE_SHADOW == 32 + 4 + 16 + 16
    assert rf.shade_bytes(10, 6, 7) == 10 * (100 + 96) + 6 * 84 + 7 * 68
    assert rf.shade_bytes(0, 0, 0) == 0
