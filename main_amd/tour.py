"""Cell tours for the locality sort (sort_hilbert_mod.f95:57-62).

The reference reads the visiting order of the cells from hilbert.txt, a file
that only exists for a 64x64x32 box (collective_code/hilbert.txt).  That file
is a 3-D Hilbert curve: four 32^3 blocks, each the order-5 curve below under a
cube symmetry, laid out in a U over x-y.  The order-k curve is self-similar:
its 8 octants are visited in an order that depends on k mod 3, and octant j
holds the order-(k-1) curve under a fixed symmetry (_LEVELS).  This module
regenerates that file exactly for 64x64x32 and, for any other box, orders the
cells along the same curve on the enclosing power-of-two cube.

A tour is returned as 1-based cell numbers  x + Lx*y + Lx*Ly*z + 1  in visiting
order (the "newvol" array the reference builds).
"""
from __future__ import annotations

import numpy as np

# symmetry = (axis permutation, flips): out[:, d] = in[:, perm[d]], then
# out[:, d] = n-1-out[:, d] where flips[d]
_LEVELS = {
    1: ([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)],
        [((0, 1, 2), (0, 0, 0)), ((2, 0, 1), (0, 0, 0)), ((2, 0, 1), (0, 0, 0)), ((1, 2, 0), (1, 1, 0)),
         ((1, 2, 0), (1, 1, 0)), ((2, 0, 1), (0, 1, 1)), ((2, 0, 1), (0, 1, 1)), ((0, 1, 2), (1, 0, 1))]),
    2: ([(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), (0, 1, 0)],
        [((0, 1, 2), (0, 0, 0)), ((2, 0, 1), (0, 0, 0)), ((2, 0, 1), (0, 0, 0)), ((1, 2, 0), (1, 0, 1)),
         ((1, 2, 0), (1, 0, 1)), ((2, 0, 1), (1, 1, 0)), ((2, 0, 1), (1, 1, 0)), ((0, 1, 2), (0, 1, 1))]),
    0: ([(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), (1, 0, 1), (1, 1, 1), (1, 1, 0), (1, 0, 0)],
        [((0, 1, 2), (0, 0, 0)), ((2, 0, 1), (0, 0, 0)), ((2, 0, 1), (0, 0, 0)), ((1, 2, 0), (0, 1, 1)),
         ((1, 2, 0), (0, 1, 1)), ((2, 0, 1), (1, 0, 1)), ((2, 0, 1), (1, 0, 1)), ((0, 1, 2), (1, 1, 0))]),
}
# blocks of the 64x64x32 file: (origin, symmetry of the order-5 curve)
_BLOCKS_64_64_32 = [((0, 0, 0), ((0, 1, 2), (0, 0, 0))), ((0, 32, 0), ((2, 0, 1), (0, 0, 0))),
                    ((32, 32, 0), ((2, 0, 1), (1, 0, 1))), ((32, 0, 0), ((0, 1, 2), (1, 1, 0)))]


def _apply(sym, pts: np.ndarray, n: int) -> np.ndarray:
    perm, flips = sym
    out = pts[:, list(perm)].copy()
    for d in range(3):
        if flips[d]:
            out[:, d] = n - 1 - out[:, d]
    return out


def hilbert_curve(k: int) -> np.ndarray:
    """(8**k, 3) int32 cell coordinates of the order-k curve on a 2^k cube."""
    cur = np.zeros((1, 3), np.int32)
    for level in range(1, k + 1):
        order, syms = _LEVELS[level % 3]
        n = 1 << (level - 1)
        parts = [_apply(syms[j], cur, n) + np.asarray(order[j], np.int32) * n for j in range(8)]
        cur = np.concatenate(parts)
    return cur


def reference_tour_coords_64x64x32() -> np.ndarray:
    """The 131072 (x, y, z) rows of collective_code/hilbert.txt, regenerated."""
    c5 = hilbert_curve(5)
    return np.concatenate([_apply(sym, c5, 32) + np.asarray(o, np.int32) for o, sym in _BLOCKS_64_64_32])


def tour_coords(Lx: int, Ly: int, Lz: int) -> np.ndarray:
    """(Lx*Ly*Lz, 3) coordinates of the cells in visiting order."""
    if (Lx, Ly, Lz) == (64, 64, 32):
        return reference_tour_coords_64x64x32()
    k = max(1, int(np.ceil(np.log2(max(Lx, Ly, Lz)))))
    c = hilbert_curve(k)
    keep = (c[:, 0] < Lx) & (c[:, 1] < Ly) & (c[:, 2] < Lz)
    return c[keep]


def cell_tour(Lx: int, Ly: int, Lz: int) -> np.ndarray:
    """1-based cell numbers in visiting order (sort_hilbert_mod.f95:60)."""
    c = tour_coords(Lx, Ly, Lz).astype(np.int64)
    return c[:, 0] + Lx * c[:, 1] + Lx * Ly * c[:, 2] + 1


def write_hilbert_txt(path: str, Lx: int, Ly: int, Lz: int) -> None:
    """Write the tour in the "x y z" per-line format sort_hilbert reads."""
    c = tour_coords(Lx, Ly, Lz)
    np.savetxt(path, c, fmt="%d", delimiter="\t")
