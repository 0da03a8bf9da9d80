"""Scene generation for the FLIP workloads (fluid masks + seeded particles).

Particles are placed with Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed by
(seed, 0) with counter (cell linear index, slot, stream, 0). The device
seeding kernel (k_seed in csrc/particles.hip) computes the same integers, so
a scene seeded on the GPU equals the one built here bit for bit.

Jitter: u = (x >> 8) * 2^-24 in [0, 1); pos = f32(i) + f32(u) rounded to
nearest, then clamped to the largest float below i + 1 so floor(pos) == i.

The reference seeds with glibc random() and rejects samples with phi > 0
(flip_simulation.cc:162-182); the workloads here specify masks directly and
place exactly `ppc` particles per fluid cell, as the benchmark scenes state.
"""
from __future__ import annotations

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Vectorised Philox4x32 with 10 rounds. Counters are uint32-valued arrays."""
    c0 = np.asarray(c0, dtype=np.uint64) & _MASK
    c1 = np.asarray(c1, dtype=np.uint64) & _MASK
    c2 = np.asarray(c2, dtype=np.uint64) & _MASK
    c3 = np.asarray(c3, dtype=np.uint64) & _MASK
    k0 &= 0xFFFFFFFF
    k1 &= 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + _W0) & 0xFFFFFFFF
            k1 = (k1 + _W1) & 0xFFFFFFFF
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0)), lo1, (hi0 ^ c3 ^ np.uint64(k1)), lo0
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def _jitter(i: np.ndarray, x: np.ndarray) -> np.ndarray:
    u = (x >> np.uint32(8)).astype(np.float32) * np.float32(2.0 ** -24)
    base = i.astype(np.float32)
    p = base + u
    top = np.nextafter((i + 1).astype(np.float32), np.float32(0))
    return np.minimum(p, top)


def seed_particles(mask: np.ndarray, ppc: int, seed: int):
    """Positions for `ppc` jittered particles in every cell where mask != 0.

    Cells are visited in x-major (i, j, k) order and slots in order, matching
    the device kernel. Returns float32 arrays pos (n, 3), vel (n, 3) (zeros).
    """
    nx, ny, nz = mask.shape
    cells = np.flatnonzero(mask.reshape(-1)).astype(np.int64)
    ci = cells // (ny * nz)
    cj = (cells // nz) % ny
    ck = cells % nz
    lin = np.repeat(cells, ppc)
    slot = np.tile(np.arange(ppc, dtype=np.int64), cells.size)
    x0, x1, x2, _ = philox4x32_10(lin, slot, 0, 0, seed, 0)
    pos = np.empty((lin.size, 3), dtype=np.float32)
    pos[:, 0] = _jitter(np.repeat(ci, ppc), x0)
    pos[:, 1] = _jitter(np.repeat(cj, ppc), x1)
    pos[:, 2] = _jitter(np.repeat(ck, ppc), x2)
    vel = np.zeros_like(pos)
    return pos, vel


def cell_centres(n):
    nx, ny, nz = n
    i = np.arange(nx, dtype=np.float64)[:, None, None] + 0.5
    j = np.arange(ny, dtype=np.float64)[None, :, None] + 0.5
    k = np.arange(nz, dtype=np.float64)[None, None, :] + 0.5
    return i, j, k


def sphere_drop_mask(n, pool_height, centre, radius) -> np.ndarray:
    """Pool below `pool_height` plus a sphere, tested at cell centres."""
    i, j, k = cell_centres(n)
    pool = np.broadcast_to(j < pool_height, tuple(n))
    d2 = (i - centre[0]) ** 2 + (j - centre[1]) ** 2 + (k - centre[2]) ** 2
    return (pool | (d2 < radius * radius)).astype(np.uint8)


def box_mask(n, lo, hi) -> np.ndarray:
    """Cells with lo <= index < hi on every axis."""
    m = np.zeros(tuple(n), dtype=np.uint8)
    m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1
    return m


def dam_break_mask(n, x_end, y_end) -> np.ndarray:
    return box_mask(n, (0, 0, 0), (x_end, y_end, n[2]))


def multi_dam_mask(n, pool_height, block) -> np.ndarray:
    """Pool plus four blocks of size `block` in the bottom corners (x, z)."""
    m = box_mask(n, (0, 0, 0), (n[0], pool_height, n[2]))
    bx, by, bz = block
    for x0 in (0, n[0] - bx):
        for z0 in (0, n[2] - bz):
            m |= box_mask(n, (x0, 0, z0), (x0 + bx, by, z0 + bz))
    return m


# The workloads, as (dims, mask builder, seed).
WORKLOADS = {
    # 64^3 sphere drop: pool j < 16, sphere r = 10 at (32, 46, 32).
    "sphere_drop_64": ((64, 64, 64), lambda n: sphere_drop_mask(n, 16, (32, 46, 32), 10), 1),
    # 128^3 dam-break projection: fluid x < 48, y < 80.
    "dam_break_128": ((128, 128, 128), lambda n: dam_break_mask(n, 48, 80), 2),
    # 256^3 dam-break FLIP step: fluid x < 96, y < 160.
    "dam_break_256": ((256, 256, 256), lambda n: dam_break_mask(n, 96, 160), 3),
    # 128^3 sphere drop scaled 2x: pool j < 32, sphere r = 20 at (64, 92, 64).
    "sphere_drop_128": ((128, 128, 128), lambda n: sphere_drop_mask(n, 32, (64, 92, 64), 20), 4),
    # 512^3 multi-dam: pool j < 32 plus four 96x192x96 corner blocks.
    "multi_dam_512": ((512, 512, 512), lambda n: multi_dam_mask(n, 32, (96, 192, 96)), 5),
}


def normal_faces(n, seed: int) -> np.ndarray:
    """I.i.d. N(0,1) float32 face velocities over the box [-1, n] (Box-Muller).

    Counter (face linear index in the box, component, stream 1, 0).
    """
    nx, ny, nz = n
    total = (nx + 2) * (ny + 2) * (nz + 2)
    lin = np.repeat(np.arange(total, dtype=np.int64), 3)
    comp = np.tile(np.arange(3, dtype=np.int64), total)
    x0, x1, _, _ = philox4x32_10(lin, comp, 1, 0, seed, 0)
    u1 = ((x0 >> np.uint32(8)).astype(np.float64) + 1.0) * 2.0 ** -24
    u2 = (x1 >> np.uint32(8)).astype(np.float64) * 2.0 ** -24
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return g.astype(np.float32).reshape(nx + 2, ny + 2, nz + 2, 3)
