"""The device tree build's partition rule, restated on the CPU and checked
against the oracle's sequential build (which follows pkdUpperPart,
pkd.c:1109-1157, and pkdActiveTypeOrder, pkd.c:1304-1355).

The device build partitions every cell of a level at once: a Hoare scan
swaps the k-th lower-part position holding an upper particle (counted up
from the cell's first position) with the k-th upper-part position holding a
lower particle (counted down from its last), so the destination of every
misplaced particle follows from two prefix counts.  This model applies that
rule level by level, exactly as the kernels do (flags, exclusive prefix sums,
ranks, swaps), and must give the oracle's permutation and node ranges."""
import numpy as np
import pytest

from oracle import oracle as O


def pair_partition(pos, lo, hi, lower):
    """One cell [lo, hi]: lower[k] says whether pos[k] belongs below."""
    F = np.concatenate([[0], np.cumsum(lower[lo:hi + 1])])   # prefix counts
    s = lo + int(F[-1])
    out = pos.copy()
    kL = {}
    for i in range(lo, s):                 # lower-part positions with an upper particle
        if not lower[i]:
            kL[(i - lo) - int(F[i - lo])] = i
    for i in range(s, hi + 1):             # upper-part positions with a lower particle
        if lower[i]:
            k = int(F[-1] - F[i - lo + 1])
            j = kL[k]
            out[[i, j]] = pos[[j, i]]
    return out, s


def build_model(r, bits, nBucket):
    n = len(r)
    perm = np.arange(n)
    for mask in (3, 2):                    # msrBuildTree's active-type orders
        perm, _ = pair_partition(perm, 0, n - 1, (bits[perm] & mask) != 0)
    cells = [(0, n - 1)]                   # creation order, one level at a time
    kids = {}
    level = [0]
    while level:
        nxt = []
        lower = np.zeros(n, bool)
        splits = []
        for c in level:
            lo, hi = cells[c]
            x = r[perm[lo:hi + 1]]
            bmin, bmax = x.min(0), x.max(0)
            if hi - lo + 1 > nBucket and np.any(bmax > bmin):
                d = 0
                for k in (1, 2):
                    if bmax[k] - bmin[k] > bmax[d] - bmin[d]:
                        d = k
                f = 0.5 * (bmin[d] + bmax[d])
                lower[lo:hi + 1] = r[perm[lo:hi + 1], d] < f
                splits.append(c)
        for c in splits:                   # all cells of the level, from one flag array
            lo, hi = cells[c]
            perm, s = pair_partition(perm, lo, hi, lower)
            assert lo < s <= hi
            kids[c] = (len(cells), len(cells) + 1)
            cells += [(lo, s - 1), (s, hi)]
            nxt += list(kids[c])
        level = nxt
    order = []                              # preorder: lower subtree, then upper
    stack = [0]
    while stack:
        c = stack.pop()
        order.append(c)
        if c in kids:
            stack += [kids[c][1], kids[c][0]]
    return perm, np.array([cells[c] for c in order])


@pytest.mark.parametrize("seed", range(6))
def test_level_synchronous_pairing_equals_sequential_hoare(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(200, 1500))
    # coarse lattice coordinates: many particles sit exactly on split planes
    r = rng.integers(0, 9, size=(n, 3)).astype(float) * 0.125 + rng.normal(0, 1e-3, (n, 3)) * (seed % 2)
    r = np.unique(r, axis=0)
    rng.shuffle(r)
    n = len(r)
    bits = rng.choice(np.array([2, 3], np.uint32), n)
    nb = int(rng.choice([4, 8, 16]))
    T = O.Tree(r, np.full(n, 1.0 / n), np.full(n, 0.01), iactive=bits, nBucket=nb)
    perm, ranges = build_model(r, bits, nb)
    oi, _ = T.nodes()
    assert np.array_equal(perm, T.perm())
    assert np.array_equal(ranges, oi[:, 1:3])
