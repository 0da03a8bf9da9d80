"""ctypes wrapper of the CPU oracle (oracle/liboracle.so, FP64 build liboracle64.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The oracle restates the
reference's algorithm; see flip_oracle.cpp for what each function follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LIB64_PATH = os.path.join(HERE, "liboracle64.so")
# Timed CPU-baseline builds, made on the machine that runs them
# (`make baseline`, -march=native): same source, -O3 (bit-identical to the
# checker) and -Ofast (the reference's flags).
VARIANT_PATHS = {"o3": os.path.join(HERE, "liboracle_o3.so"),
                 "ofast": os.path.join(HERE, "liboracle_ofast.so")}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_libs = {}


def build_baseline():
    """Builds the -march=native baseline variants here (make baseline)."""
    subprocess.run(["make", "-s", "-C", HERE, "baseline"], check=True)


def load(fp64=False, variant=None):
    """The FP32 restatement (liboracle.so), or with fp64 the same source built
    with double grids, particles and scalars (liboracle64.so); variant "o3"
    or "ofast" loads a baseline build (FP32)."""
    key = variant or fp64
    if key not in _libs:
        path = VARIANT_PATHS[variant] if variant else (LIB64_PATH if fp64 else LIB_PATH)
        if not os.path.exists(path):
            build_baseline() if variant else build()
        lib = C.CDLL(path)
        P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        F = C.c_double if fp64 else C.c_float
        sig = {
            "or_create": ([I, I, I], P), "or_destroy": ([P], None),
            "or_set_params": ([P, F, F, F, F, F], None), "or_set_box_boundary": ([P, I], None),
            "or_add_solid_box": ([P, P, P], None),
            "or_set_particles": ([P, I64, P, P, P], None), "or_num_particles": ([P], I64),
            "or_get_particles": ([P, P, P, P, P], None), "or_cfl": ([P], F),
            "or_move_particles": ([P, F], None), "or_clear": ([P], None), "or_p2g": ([P], None),
            "or_apply_bc": ([P], None), "or_save_velocities": ([P], None),
            "or_delete_outside": ([P], None), "or_add_gravity": ([P, F], None),
            "or_divergence": ([P], None),
            "or_solve": ([P, I, I, D, I, C.POINTER(I), C.POINTER(D)], I),
            "or_solve_pc": ([P, I, I, D, I, I, P, C.POINTER(I), C.POINTER(D)], I),
            "or_ic_factor": ([P, P, P, P, P], I),
            "or_solver_sizes": ([P, C.POINTER(I), C.POINTER(I)], I),
            "or_residual_history": ([P, P, I], I), "or_solution": ([P, P, I], I),
            "or_scalar_history": ([P, I, P, I], I), "or_set_threads": ([P, I], None),
            "or_pressure_update": ([P], None), "or_velocity_update": ([P], I),
            "or_extrapolate": ([P], None), "or_g2p": ([P], None),
            "or_run_one_step": ([P, F, I, I, D, I, C.POINTER(I)], I),
            "or_compute_dt": ([P, I, F, C.POINTER(F)], F),
            "or_get_vec": ([P, I, P, P], None), "or_set_vec": ([P, I, P, P], None),
            "or_get_scalar": ([P, I, P], None), "or_set_scalar": ([P, I, P], None),
            "or_get_int": ([P, I, P], None), "or_set_marker": ([P, P], None),
            "or_cell_counts": ([P, P], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _libs[key] = lib
    return _libs[key]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The reference's one-partition FLIP step on dense CPU grids."""

    def __init__(self, dims, fp64=False, variant=None, threads=None):
        self.lib = load(fp64, variant)
        self.real = np.float64 if fp64 else np.float32
        self.creal = C.c_double if fp64 else C.c_float
        self.dims = tuple(int(d) for d in dims)
        self.box = tuple(d + 2 for d in self.dims)
        self.h = self.lib.or_create(*self.dims)
        if threads is not None:
            self.set_threads(threads)

    def set_threads(self, threads):
        """Threads of the solver's parallel loops (default: OpenMP's count)."""
        self.lib.or_set_threads(self.h, int(threads))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def set_params(self, gravity_scaled, force, flip):
        self.lib.or_set_params(self.h, gravity_scaled, force[0], force[1], force[2], flip)

    def set_box_boundary(self, tank=False):
        self.lib.or_set_box_boundary(self.h, int(tank))

    def add_solid_box(self, lo, hi):
        self.lib.or_add_solid_box(self.h, _ptr(np.array(lo, np.int32)), _ptr(np.array(hi, np.int32)))

    def set_particles(self, pos, vel, gv=None):
        pos = np.ascontiguousarray(pos, self.real)
        vel = np.ascontiguousarray(vel, self.real)
        gv = vel if gv is None else np.ascontiguousarray(gv, self.real)
        self.lib.or_set_particles(self.h, pos.shape[0], _ptr(pos), _ptr(vel), _ptr(gv))

    def num_particles(self):
        return int(self.lib.or_num_particles(self.h))

    def get_particles(self):
        n = self.num_particles()
        pos, vel, gv = (np.empty((n, 3), self.real) for _ in range(3))
        cell = np.empty((n, 3), np.int32)
        self.lib.or_get_particles(self.h, _ptr(pos), _ptr(vel), _ptr(gv), _ptr(cell))
        return pos, vel, gv, cell

    def cfl(self):
        return float(self.lib.or_cfl(self.h))

    def compute_dt(self, cfl, frame_step, step):
        s = self.creal(step)
        dt = self.lib.or_compute_dt(self.h, int(cfl), float(frame_step), C.byref(s))
        return float(dt), s.value

    def move_particles(self, dt):
        self.lib.or_move_particles(self.h, float(dt))

    def clear(self):
        self.lib.or_clear(self.h)

    def p2g(self):
        self.lib.or_p2g(self.h)

    def apply_bc(self):
        self.lib.or_apply_bc(self.h)

    def save_velocities(self):
        self.lib.or_save_velocities(self.h)

    def delete_outside(self):
        self.lib.or_delete_outside(self.h)

    def add_gravity(self, dt):
        self.lib.or_add_gravity(self.h, float(dt))

    def divergence(self):
        self.lib.or_divergence(self.h)

    def solve(self, precision=0, stop_mode=0, threshold=None, max_iter=1500, precond=0,
              tiles=None):
        """PCG on the current divergence; precond 0 Jacobi, 1 block IC(0) over
        `tiles` (partitions per axis; None: one). Returns (rc, iterations, rr)."""
        if threshold is None:
            threshold = float(np.float32(np.float32(np.prod(self.dims)) * np.float32(0.5e-8)))
        it, rr = C.c_int(), C.c_double()
        t = None if tiles is None else np.ascontiguousarray(tiles, np.int32)
        rc = self.lib.or_solve_pc(self.h, precision, stop_mode, threshold, max_iter, precond,
                                  _ptr(t), C.byref(it), C.byref(rr))
        return rc, it.value, rr.value

    def ic_factor(self, tiles=None, r=None):
        """Block IC(0) of the current system: (L diagonal in row order, and
        z = M^-1 r when r is given), or None when the build fails."""
        cap = int(np.prod(self.dims))
        diag = np.zeros(cap, np.float64)
        t = None if tiles is None else np.ascontiguousarray(tiles, np.int32)
        rv = None if r is None else np.ascontiguousarray(r, np.float64)
        z = None if r is None else np.zeros(cap, np.float64)
        m = self.lib.or_ic_factor(self.h, _ptr(t), _ptr(diag), _ptr(rv), _ptr(z))
        if m < 0:
            return None
        return diag[:m].copy(), (None if z is None else z[:m].copy())

    def solver_sizes(self):
        a, b = C.c_int(), C.c_int()
        self.lib.or_solver_sizes(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def residual_history(self):
        out = np.empty(4096, np.float64)
        m = self.lib.or_residual_history(self.h, _ptr(out), out.size)
        return out[: min(m, out.size)].copy()

    def scalar_history(self):
        """Scalars of the last solve: dict of rr and rz (after prepare, then
        after each iteration) and pAp, alpha, beta (one per iteration)."""
        out = {}
        for w, name in enumerate(("rr", "pAp", "alpha", "rz", "beta")):
            buf = np.empty(20000, np.float64)
            m = self.lib.or_scalar_history(self.h, w, _ptr(buf), buf.size)
            out[name] = buf[: min(m, buf.size)].copy()
        return out

    def solution(self):
        n, _ = self.solver_sizes()
        out = np.empty(max(n, 1), np.float64)
        m = self.lib.or_solution(self.h, _ptr(out), out.size)
        return out[:m].copy()

    def pressure_update(self):
        self.lib.or_pressure_update(self.h)

    def velocity_update(self):
        return int(self.lib.or_velocity_update(self.h))

    def extrapolate(self):
        self.lib.or_extrapolate(self.h)

    def g2p(self):
        self.lib.or_g2p(self.h)

    def run_one_step(self, dt, precision=0, stop_mode=0, threshold=None, max_iter=1500):
        if threshold is None:
            threshold = float(np.float32(np.float32(np.prod(self.dims)) * np.float32(0.5e-8)))
        it = C.c_int()
        rc = self.lib.or_run_one_step(self.h, float(dt), precision, stop_mode, threshold,
                                      max_iter, C.byref(it))
        return rc, it.value

    def get_vec(self, which=0):
        out = np.empty(self.box + (3,), self.real)
        act = np.empty(self.box, np.uint8)
        self.lib.or_get_vec(self.h, which, _ptr(out), _ptr(act))
        return out, act

    def set_vec(self, which, values, active):
        v = np.ascontiguousarray(values, self.real)
        a = np.ascontiguousarray(active, np.uint8)
        self.lib.or_set_vec(self.h, which, _ptr(v), _ptr(a))

    def get_scalar(self, which):
        out = np.empty(self.box, self.real)
        self.lib.or_get_scalar(self.h, which, _ptr(out))
        return out

    def set_scalar(self, which, values):
        self.lib.or_set_scalar(self.h, which, _ptr(np.ascontiguousarray(values, self.real)))

    def get_int(self, which):
        out = np.empty(self.box, np.int32)
        self.lib.or_get_int(self.h, which, _ptr(out))
        return out

    def set_marker(self, marker):
        self.lib.or_set_marker(self.h, _ptr(np.ascontiguousarray(marker, np.int32)))

    def cell_counts(self):
        out = np.empty(self.box, np.int32)
        self.lib.or_cell_counts(self.h, _ptr(out))
        return out


# ---------------------------------------------------------------- sources
# Numpy restatements (no C part): the water sources of the reference with
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11; 10 rounds, Weyl key
# schedule) in place of glibc random().

_PH_M0, _PH_M1 = 0xD2511F53, 0xCD9E8D57
_PH_W0, _PH_W1 = 0x9E3779B9, 0xBB67AE85
_U32 = 0xFFFFFFFF


def philox(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 of uint32 counter arrays under the key (k0, k1)."""
    c = [np.asarray(x, np.uint64) & _U32 for x in (c0, c1, c2, c3)]
    c = np.broadcast_arrays(*c)
    c = [x.copy() for x in c]
    k0, k1 = int(k0) & _U32, int(k1) & _U32
    for r in range(10):
        if r:
            k0 = (k0 + _PH_W0) & _U32
            k1 = (k1 + _PH_W1) & _U32
        p0 = np.uint64(_PH_M0) * c[0]
        p1 = np.uint64(_PH_M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(_U32)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(_U32)
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return [x.astype(np.uint32) for x in c]


def source_particles(dims, lo, hi, velocity, ppc, seed, stream):
    """AddContributionFromSources (flip_simulation.cc:1206-1266), particles
    only, for one box as Source::get_box returns it (end already reduced by
    1). Cells from int(lo) to int(hi) (truncation, as the Coord conversion),
    clipped to the domain, in x-major order; each makes up to 32 attempts,
    attempt a giving the point cell + u with u = (x >> 8) * 2^-24 from
    philox((lin, a, 2, 0), (seed, stream)), lin the cell's x-major index;
    points inside the box (bounds inclusive, Cube::IsInside,
    primitives.cc:57-64) are kept, the first ppc of them. Returns positions
    (n, 3) float32 in emission order and velocities (n, 3)."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    c0 = [max(int(lo[e]), 0) for e in range(3)]
    c1 = [min(int(hi[e]), dims[e] - 1) for e in range(3)]
    if any(c1[e] < c0[e] for e in range(3)):
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)
    ax = [np.arange(c0[e], c1[e] + 1) for e in range(3)]
    cells = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    lin = (cells[:, 0].astype(np.int64) * dims[1] + cells[:, 1]) * dims[2] + cells[:, 2]
    att = np.arange(32, dtype=np.int64)
    x = philox(lin[:, None], att[None, :], 2, 0, seed, stream)
    pos = np.empty(cells.shape[:1] + (32, 3), np.float32)
    for e in range(3):
        u = (x[e] >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
        pos[..., e] = cells[:, e, None].astype(np.float32) + u
    inside = np.all((pos >= lo) & (pos <= hi), axis=-1)
    keep = inside & (np.cumsum(inside, axis=1) <= ppc)
    out = pos[keep]
    vel = np.tile(np.asarray(velocity, np.float32), (out.shape[0], 1))
    return out, vel


def inside_boxes(pos, boxes):
    """Mask of the particles that DeleteParticlesFromRegions (particles.cc:
    222-271) removes: position inside any box (bounds inclusive)."""
    pos = np.asarray(pos, np.float32)
    m = np.zeros(pos.shape[0], bool)
    for b in np.asarray(boxes, np.float32).reshape(-1, 6):
        m |= np.all((pos >= b[:3]) & (pos <= b[3:]), axis=1)
    return m
