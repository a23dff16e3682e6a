"""The rule by which the LBVH's bottom-up pass (k_lbvh, lr_build.hip) keeps a climb inside its
work-group: a node's parent lies inside the group's leaves exactly when a min-scan of the group's
adjacent-key deltas says so. Checked against the parent ranges the radix tree itself gives, for
every node of trees over random, duplicated and heavily equal keys, with groups of several sizes.

The tree is the binary radix tree of the sorted keys with Karras's delta (trace_build.wgsl:52-58:
the common prefix of adjacent codes, 32 + that of their indices when the codes are equal): a
node's parent splits at its right end when delta(r) > delta(l - 1), else at l - 1. Its true
parent range comes from the parent's split p: for a left child [l, p] it ends at the first
k > p with delta(k) < delta(p) (or n - 1), for a right child [p + 1, r] it starts after the last
k < p with delta(k) < delta(p) (or 0); and the parent found so must have its split at p (checked).
"""
import numpy as np
import pytest


def deltas(codes):
    n = len(codes)
    d = np.full(n + 1, -1, np.int64)            # d[k + 1] = delta(k), -1 off the ends
    for k in range(n - 1):
        a, b = int(codes[k]), int(codes[k + 1])
        d[k + 1] = 32 + (32 - (k ^ (k + 1)).bit_length()) if a == b else 32 - (a ^ b).bit_length()
    return d


def check(codes, group):
    n = len(codes)
    d = deltas(codes)
    D = lambda k: int(d[k + 1]) if -1 <= k < n else -1
    seen, inside_count, total = set(), 0, 0
    for j in range(n):
        l = r = j
        while not (l == 0 and r == n - 1):
            left = D(r) > D(l - 1)
            p = r if left else l - 1
            if left:
                R = next((k for k in range(p + 1, n) if D(k) < D(p)), n - 1)
                L = l
            else:
                L = next((k + 1 for k in range(p - 1, -2, -1) if D(k) < D(p)), 0)
                R = r
            # the parent [L, R] splits at p: its smallest inner delta is there
            assert min(D(k) for k in range(L, R)) == D(p)
            if (l, r) in seen:
                break
            seen.add((l, r))
            B0 = (j // group) * group
            if l >= B0 and r <= B0 + group - 1:
                # the kernel's test: d of the group's keys and the one before (off the end: -1)
                if left:
                    ins = p <= B0 + group - 2 and min(D(k) for k in range(p + 1, B0 + group)) < D(p)
                else:
                    ins = p >= B0 and min(D(k) for k in range(B0 - 1, p)) < D(p)
                assert ins == (L >= B0 and R <= B0 + group - 1), (l, r, p, L, R, B0)
                inside_count += ins
                total += 1
            l, r = L, R
    assert len(seen) == 2 * n - 2           # every node but the root, once
    return inside_count, total


@pytest.mark.parametrize("group", [4, 16, 512])
@pytest.mark.parametrize("kind", ["random", "duplicated", "few"])
def test_group_rule_equals_true_containment(group, kind):
    rng = np.random.default_rng(7)
    for n in [2, 3, 5, 17, 100, 1025]:
        if kind == "random":
            c = rng.integers(0, 1 << 30, n)
        elif kind == "duplicated":
            c = rng.integers(0, 1 << 30, n)
            c[::3] = c[0]
        else:
            c = rng.integers(0, 8, n)
        inside, total = check(np.sort(c).astype(np.uint32), group)
        if n > 2 * group:
            assert 0 < inside < total
