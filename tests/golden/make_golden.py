"""Writes tests/golden/projection_systems.npz: pressure systems with answers
computed here, independently of the oracle and of the library.

For each system: a label grid (box format [-1, n], ring SOLID, random FLUID,
AIR and interior SOLID cells, every FLUID component touching AIR), FP32 face
velocities that vanish on the domain boundary (so ApplyBoundaryConditions
leaves them unchanged), and
  b      the divergence on FLUID cells in FP32, accumulated as the reference
         does (0 + du + dv + dw, flip_simulation.cc:715-736);
  x      the FP64 dense direct solution of A x = b (scipy), A the Laplacian of
         solver.cc:197-233 (diagonal = FLUID or AIR neighbours, -1 per FLUID
         neighbour), rows in lexicographic order (solver.cc:114-151);
  ic     the IC(0) diagonal of A in FP64 (one tile), from the textbook
         sparse recurrence on the dense pattern, and ic_tiles the same with
         couplings between tiles removed (block IC over `tiles`).
Run from the repository root: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg
from scipy import ndimage

HERE = os.path.dirname(os.path.abspath(__file__))
SYSTEMS = [((8, 8, 8), (1, 1, 1), 41), ((16, 8, 24), (2, 1, 3), 42), ((16, 16, 16), (2, 2, 2), 43)]


def labels(n, rng):
    box = tuple(d + 2 for d in n)
    u = rng.random(n)
    inner = np.where(u < 0.8, 1, np.where(u < 0.85, 2, 0))
    marker = np.full(box, 2, np.int8)
    marker[1:-1, 1:-1, 1:-1] = inner
    lab, k = ndimage.label(marker == 1)
    air = ndimage.binary_dilation(marker == 0)
    for comp in range(1, k + 1):
        if not np.any(air[lab == comp]):
            marker[lab == comp] = 0
    return marker


def faces(n, rng):
    box = tuple(d + 2 for d in n)
    f = rng.standard_normal(box + (3,)).astype(np.float32)
    for d in range(3):
        sl = [slice(None)] * 3
        for idx in (1, n[d] + 1):  # faces between the ring and the domain
            sl[d] = idx
            f[tuple(sl) + (d,)] = 0.0
    return f


def divergence(marker, f):
    b = []
    for c in np.argwhere(marker == 1):
        i, j, k = c
        acc = np.float32(0.0)
        acc = acc + (f[i + 1, j, k, 0] - f[i, j, k, 0])
        acc = acc + (f[i, j + 1, k, 1] - f[i, j, k, 1])
        acc = acc + (f[i, j, k + 1, 2] - f[i, j, k, 2])
        b.append(acc)
    return np.array(b, np.float32)


def laplacian(marker):
    fl = np.argwhere(marker == 1)
    index = {tuple(c): r for r, c in enumerate(fl)}
    A = np.zeros((len(fl), len(fl)))
    for r, c in enumerate(fl):
        for d in range(3):
            for s in (-1, 1):
                nb = c.copy()
                nb[d] += s
                m = marker[tuple(nb)]
                if m != 2:
                    A[r, r] += 1
                if m == 1:
                    A[r, index[tuple(nb)]] = -1
    return A, fl


def ic0_diag(A, fl, n, tiles):
    """IC(0) by the column-oriented recurrence restricted to A's pattern."""
    w = np.array(n) // np.array(tiles)
    tile = [tuple((c - 1) // w) for c in fl]
    m = A.shape[0]
    cols = [dict() for _ in range(m)]  # lower part, row -> {col: value}
    for i in range(m):
        for j in np.nonzero(A[i])[0]:
            if j <= i and tile[i] == tile[j]:
                cols[i][j] = A[i, j]
    for k in range(m):
        cols[k][k] = np.sqrt(cols[k][k])
        # Rows with an entry in column k are the upper neighbours of k.
        nbrs = [j for j in np.nonzero(A[k])[0] if j > k and tile[j] == tile[k]]
        for i in nbrs:
            cols[i][k] = cols[i][k] / cols[k][k]
        for j in nbrs:
            for i in [i for i in np.nonzero(A[j])[0] if i >= j and tile[i] == tile[j]]:
                if k in cols[i]:
                    cols[i][j] = cols[i][j] - cols[i][k] * cols[j][k]
    return np.array([cols[i][i] for i in range(m)])


def main():
    out = {}
    for s, (n, tiles, seed) in enumerate(SYSTEMS):
        rng = np.random.default_rng(seed)
        marker = labels(n, rng)
        f = faces(n, rng)
        b = divergence(marker, f)
        A, fl = laplacian(marker)
        x = scipy.linalg.solve(A, b.astype(np.float64), assume_a="pos")
        out[f"dims_{s}"] = np.array(n, np.int32)
        out[f"tiles_{s}"] = np.array(tiles, np.int32)
        out[f"marker_{s}"] = marker
        out[f"faces_{s}"] = f
        out[f"b_{s}"] = b
        out[f"x_{s}"] = x
        out[f"ic_{s}"] = ic0_diag(A, fl, n, (1, 1, 1))
        out[f"ic_tiles_{s}"] = ic0_diag(A, fl, n, tiles)
        print(f"system {s}: dims {n}, {len(b)} rows, tiles {tiles}")
    np.savez_compressed(os.path.join(HERE, "projection_systems.npz"), **out)


if __name__ == "__main__":
    main()
