"""Algorithmic bytes of the hot kernels: one formula per kernel, the bytes the algorithm must move
between the compute units and memory, whatever the caches do. A kernel's roofline figure is these
bytes over its measured time, against the HBM peak.

Walks (k_trace, k_extend, k_shadow, k_scene_trace): per ray its 32-byte ray read and its 16-byte
hit written, per node record fetched its size (32 bytes for the reference's binary node, 64 for
the compressed 4-wide node), per shape tested its 48-byte triangle and 4-byte aux entry. The node
and shape counts are the walks' own counters on the same tree and rays.

LBVH build (k_bounds, k_morton, k_sort_pass x 4, k_lbvh and k_lbvh_top), per primitive:
  bounds    the box read, 32
  morton    the box read, code and item written, 32 + 8
  sort      four 8-bit passes, each (code, item) read and written, 4 x 16
  lbvh      item 4 and box 32 read, leaf node 32 and aux 4 written, and per inner node (n - 1 of
            them, counted per primitive) the sibling's record read and the node written, 32 + 32

Shade (k_shade), per queue entry (one per closest-hit ray): its state read, the queue's slot id 4,
meta 8, ray 32, hit 16, throughput 16, radiance 16, generator 8 (100); the hit's triangle 48 and
material 48 (counted for every entry: an escaped ray reads neither, so this part is an upper
bound); per path that goes on, its next entry written (ray 32, slot id 4, meta 8, throughput 16,
radiance 16, generator 8: 84); per shadow ray, its entry written (ray 32, id 4, the path's place
16, radiance 16: 68). The BSDF tables and blue-noise ranks (12 KB, read by every entry) are not
counted: they stay in the caches.

PLOC build (k_ploc_init, k_ploc_round x rounds, k_ploc_small), the bounds, Morton and sort
passes before it priced as for the LBVH (bounds 32, morton 40, sort 64 per primitive):
  init      per primitive: item 4 and box 32 read, cluster box 32, aux 4, height 4 and list
            entry 4 written, 80
  rounds    per live cluster of each round: its list entry 4 and box 32 read, the next list's
            entry 4 written, 40 (live: the clusters summed over the rounds, the build's count)
  merges    n - 1 of them: the merged box 32, the two children's records 64 written, their
            heights 8 read and the merged one 4 written, 108
"""

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E

NODE_BYTES_BINARY = 32          # the reference's node (put_nodes)
NODE_BYTES_WIDE = 64            # the compressed 4-wide node
RAY_BYTES, HIT_BYTES, SHAPE_BYTES = 32, 16, 52

LBVH_BYTES = {"bounds": 32, "morton": 40, "sort": 64, "lbvh": 4 + 32 + 32 + 4 + 64}

SHADE_IN, SHADE_HIT, SHADE_NEXT, SHADE_SHADOW = 100, 96, 84, 68


def shade_bytes(entries, continued, shadow_rays):
    """the shade kernel's algorithmic bytes: entries shaded, paths that went on, shadow rays made"""
    return (SHADE_IN + SHADE_HIT) * entries + SHADE_NEXT * continued + SHADE_SHADOW * shadow_rays


def walk_bytes(rays, node_fetches, shape_tests, node_bytes):
    """a walk kernel's algorithmic bytes"""
    return (RAY_BYTES + HIT_BYTES) * rays + node_bytes * node_fetches + SHAPE_BYTES * shape_tests


def lbvh_bytes(n, with_bounds=True):
    """an LBVH build of n primitives (the bounds pass included unless the caller gives them)"""
    per = sum(v for k, v in LBVH_BYTES.items() if with_bounds or k != "bounds")
    return per * n


PLOC_INIT, PLOC_PER_LIVE, PLOC_PER_MERGE = 80, 40, 108


def ploc_bytes(n, live, with_bounds=True):
    """a PLOC build of n primitives whose rounds held `live` clusters in all"""
    pre = sum(v for k, v in LBVH_BYTES.items() if k != "lbvh" and (with_bounds or k != "bounds"))
    return (pre + PLOC_INIT) * n + PLOC_PER_LIVE * live + PLOC_PER_MERGE * max(n - 1, 0)


def roofline(alg_bytes, ms):
    """(achieved GB/s, fraction of the HBM peak) of alg_bytes moved in ms milliseconds"""
    gbs = alg_bytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return gbs, gbs / HBM_PEAK_GBS
