"""Python binding of libfsl (the MI355X FLIP projection core) over its C ABI.

The binding is a thin ctypes layer: every call goes to the HIP library
built in tree (fluid-sim-lb_amd/libfsl.so); there is no CPU fallback. If the
library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scenes  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libfsl.so")

OK = 0
ERR = {
    -1: "FSL_ERR_ARG", -2: "FSL_ERR_HIP", -3: "FSL_ERR_STATE", -4: "FSL_ERR_CAPACITY",
    -10: "FSL_ERR_ZERO_DIAGONAL", -11: "FSL_ERR_NAN", -12: "FSL_ERR_ALPHA",
    -13: "FSL_ERR_BETA", -14: "FSL_ERR_VELOCITY_UPDATE", -15: "FSL_ERR_PRECONDITIONER",
}
PCG_FP32, PCG_FP64, PCG_MIXED = 0, 1, 2
PCG_AUTO, PCG_TWO_PASS, PCG_SINGLE_REDUCTION = 0, 1, 2
STOP_ABS_RR, STOP_REL = 0, 1
PRECOND_JACOBI, PRECOND_IC = 0, 1
DOT_TREE, DOT_ROW_ORDER = 0, 1
GRID_VELOCITY, GRID_VELOCITY_UPDATE = 0, 1
GRID_DIVERGENCE, GRID_PRESSURE = 0, 1
GRID_MARKER, GRID_SOLID, GRID_ROW_INDEX, GRID_CELL_COUNT = 0, 1, 2, 3
TIMING_NAMES = ("move", "clear", "p2g", "bc_save", "gravity_bc", "divergence", "solve",
                "update", "g2p", "pass_a", "pass_b", "pass_a_n", "pass_b_n", "reduce", "steps",
                "fused", "fused_n")


class SolveParams(C.Structure):
    _fields_ = [("precision", C.c_int), ("stop_mode", C.c_int), ("threshold", C.c_double),
                ("max_iterations", C.c_int), ("refresh_every", C.c_int), ("batch", C.c_int),
                ("warm_start", C.c_int), ("variant", C.c_int), ("preconditioner", C.c_int),
                ("dot_order", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("n_rows", C.c_int), ("n_cols", C.c_int),
                ("n_fluid_leaves", C.c_int), ("rr", C.c_double),
                ("true_rel_residual", C.c_double), ("bnorm", C.c_double),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class FrameStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("pcg_iterations", C.c_int), ("dt_last", C.c_float),
                ("last_solve", SolveStats)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_[:3]}
        d["last_solve"] = self.last_solve.as_dict()
        return d


# Exported symbols and their argument types (the declarations of include/fsl.h).
_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
_I64 = C.c_int64
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
SIGNATURES = {
    "fsl_create": [C.POINTER(_P), _IP, _IP, _I, _I64],
    "fsl_destroy": [_P],
    "fsl_last_error": [_P],
    "fsl_set_params": [_P, _F, _FP, _F],
    "fsl_set_box_boundary": [_P, _I],
    "fsl_add_solid_box": [_P, _IP, _IP],
    "fsl_set_particles": [_P, _I64, _P, _P, _P],
    "fsl_seed_particles": [_P, _P, _I, C.c_uint32, _FP],
    "fsl_add_source": [_P, _FP, _FP, _FP, _I, C.c_uint32, C.c_uint32],
    "fsl_delete_particles_in_boxes": [_P, _I, _FP],
    "fsl_num_particles": [_P, C.POINTER(_I64)],
    "fsl_get_particles": [_P, _P, _P, _P],
    "fsl_cfl": [_P, _FP],
    "fsl_move_particles": [_P, _F],
    "fsl_clear": [_P],
    "fsl_p2g": [_P],
    "fsl_apply_bc": [_P],
    "fsl_save_velocities": [_P],
    "fsl_delete_outside": [_P],
    "fsl_add_gravity": [_P, _F],
    "fsl_divergence": [_P],
    "fsl_pcg_solve": [_P, C.POINTER(SolveParams), C.POINTER(SolveStats)],
    "fsl_pressure_update": [_P],
    "fsl_velocity_update": [_P],
    "fsl_extrapolate": [_P],
    "fsl_g2p": [_P],
    "fsl_step": [_P, _F, C.POINTER(SolveParams), C.POINTER(SolveStats)],
    "fsl_advance_frame": [_P, _I, _F, C.POINTER(SolveParams), C.POINTER(FrameStats)],
    "fsl_set_step_graph": [_P, _I],
    "fsl_compute_dt": [_P, _I, _F, _FP, _FP],
    "fsl_get_vec": [_P, _I, _P, _P],
    "fsl_set_vec": [_P, _I, _P, _P],
    "fsl_get_scalar": [_P, _I, _P],
    "fsl_get_scalar_noalloc": [_P, _I, _P],
    "fsl_set_scalar": [_P, _I, _P],
    "fsl_get_int": [_P, _I, _P],
    "fsl_set_marker": [_P, _P],
    "fsl_get_active_leaves": [_P, _I, _P, _I64, C.POINTER(_I64)],
    "fsl_tile_rows": [_P, C.POINTER(_I64)],
    "fsl_pcg_prepare": [_P, C.POINTER(SolveParams), _DP, _DP],
    "fsl_pcg_local1": [_P, _DP],
    "fsl_pcg_local2": [_P, _DP, _DP, _DP],
    "fsl_pcg_local3": [_P, _DP],
    "fsl_pcg_iterate": [_P, _DP, _DP, _DP, _DP, _DP],
    "fsl_pcg_finish": [_P, C.POINTER(SolveStats)],
    "fsl_set_timing": [_P, _I],
    "fsl_get_timing": [_P, _DP, _I],
    "fsl_project_forward": [_P, C.POINTER(SolveParams), C.POINTER(SolveStats)],
    "fsl_project_backward": [_P, C.POINTER(SolveParams), _P, _P, C.POINTER(SolveStats)],
    "fsl_synchronize": [_P],
    "fsl_rollout_begin": [_P, C.c_int],
    "fsl_rollout_step": [_P, C.c_float, C.POINTER(SolveParams), _P, _P, C.POINTER(SolveStats)],
    "fsl_rollout_replay": [_P, C.POINTER(SolveParams), _P, _P, _P, C.POINTER(SolveStats)],
    "fsl_rollout_backward": [_P, C.POINTER(SolveParams), _P, _P, _P, C.POINTER(SolveStats)],
    "fsl_rollout_checkpoint": [_P, C.c_int, _P, _P, _P, C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64), C.POINTER(C.c_float)],
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libfsl.so (raises OSError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"libfsl.so not built: {path} (run fluid-sim-lb_amd/build.py)")
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "fsl_last_error" else C.c_int
    _lib = lib
    return lib


class FslError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR.get(code, code)}: {msg}")
        self.code = code


def _ptr(a):
    """Address of a numpy array, or of a torch tensor's storage (device
    tensors are read and written in place by the entry points that accept
    device arrays)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return C.c_void_p(a.data_ptr())


def _is_device_tensor(a) -> bool:
    return not isinstance(a, np.ndarray) and getattr(a, "is_cuda", False)


def solve_params(precision=PCG_FP32, stop_mode=STOP_ABS_RR, threshold=None, dims=None,
                 max_iterations=1500, refresh_every=32, batch=0,
                 variant=PCG_AUTO, preconditioner=PRECOND_JACOBI,
                 dot_order=DOT_TREE) -> SolveParams:
    """Solver parameters; by default the reference's rule r.r < N^3 * 0.5e-8
    (flip_app.cc:326-328, flip_app.h:52-53). `variant` picks the iteration
    structure: PCG_AUTO (two-pass for FP32 or IC, single-reduction otherwise),
    PCG_TWO_PASS or PCG_SINGLE_REDUCTION. `preconditioner`: PRECOND_JACOBI
    (what the reference runs) or PRECOND_IC (block IC(0) per tile).
    `dot_order`: DOT_TREE, or DOT_ROW_ORDER (two-pass FP32/FP64 only: dot
    products summed sequentially in the solver's row order)."""
    if threshold is None:
        vol = int(np.prod(dims)) if dims is not None else 64 ** 3
        threshold = float(np.float32(np.float32(vol) * np.float32(0.5e-8)))
    return SolveParams(precision, stop_mode, threshold, max_iterations, refresh_every, batch, 0,
                       variant, preconditioner, dot_order)


class Context:
    """One simulation domain on one GPU (an fsl_ctx)."""

    def __init__(self, dims, tiles=(1, 1, 1), device=0, capacity=0):
        self.lib = load()
        self.dims = tuple(int(d) for d in dims)
        self.box = tuple(d + 2 for d in self.dims)
        self.force = (0.0, 9.8, 0.0)
        h = C.c_void_p()
        dv = (C.c_int * 3)(*self.dims)
        tv = (C.c_int * 3)(*tiles)
        rc = self.lib.fsl_create(C.byref(h), dv, tv, int(device), int(capacity))
        if rc != OK:
            raise FslError(rc, "fsl_create failed")
        self.h = h
        self.tiles = tuple(tiles)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fsl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            msg = self.lib.fsl_last_error(self.h)
            raise FslError(rc, msg.decode() if msg else "")
        return rc

    def call(self, name, *args):
        return self._check(getattr(self.lib, name)(self.h, *args))

    # -- set-up
    def set_params(self, gravity_scaled, force, flip):
        self.force = tuple(float(x) for x in force)
        f = (C.c_float * 3)(*force)
        self.call("fsl_set_params", float(gravity_scaled), f, float(flip))

    def set_box_boundary(self, tank=False):
        self.call("fsl_set_box_boundary", int(tank))

    def add_solid_box(self, lo, hi):
        self.call("fsl_add_solid_box", (C.c_int * 3)(*lo), (C.c_int * 3)(*hi))

    def set_particles(self, pos, vel, gv=None):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        gv = None if gv is None else np.ascontiguousarray(gv, dtype=np.float32)
        self.call("fsl_set_particles", int(pos.shape[0]), _ptr(pos), _ptr(vel), _ptr(gv))

    def seed_particles(self, mask, ppc, seed, velocity=(0.0, 0.0, 0.0)):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        assert m.shape == self.dims
        self.call("fsl_seed_particles", _ptr(m), int(ppc), int(seed), (C.c_float * 3)(*velocity))

    def add_source(self, lo, hi, velocity, ppc, seed, stream=0):
        """One evaluated source box (end already reduced by 1, as
        Source::get_box returns it); appends its particles."""
        f3 = C.c_float * 3
        self.call("fsl_add_source", f3(*lo), f3(*hi), f3(*velocity), int(ppc), int(seed),
                  int(stream))

    def delete_particles_in_boxes(self, boxes):
        """boxes: (n, 6) array of lo x, y, z, hi x, y, z (bounds inclusive)."""
        b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
        self.call("fsl_delete_particles_in_boxes", len(b),
                  b.ctypes.data_as(C.POINTER(C.c_float)))

    def num_particles(self) -> int:
        n = C.c_int64()
        self.call("fsl_num_particles", C.byref(n))
        return n.value

    def get_particles(self):
        n = self.num_particles()
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        gv = np.empty((n, 3), np.float32)
        self.call("fsl_get_particles", _ptr(pos), _ptr(vel), _ptr(gv))
        return pos, vel, gv

    # -- stages
    def cfl(self) -> float:
        x = C.c_float()
        self.call("fsl_cfl", C.byref(x))
        return x.value

    def compute_dt(self, cfl, frame_step, step):
        s = C.c_float(step)
        dt = C.c_float()
        self.call("fsl_compute_dt", int(cfl), float(frame_step), C.byref(s), C.byref(dt))
        return dt.value, s.value

    def move_particles(self, dt):
        self.call("fsl_move_particles", float(dt))

    def clear(self):
        self.call("fsl_clear")

    def p2g(self):
        self.call("fsl_p2g")

    def apply_bc(self):
        self.call("fsl_apply_bc")

    def save_velocities(self):
        self.call("fsl_save_velocities")

    def delete_outside(self):
        self.call("fsl_delete_outside")

    def add_gravity(self, dt):
        self.call("fsl_add_gravity", float(dt))

    def divergence(self):
        self.call("fsl_divergence")

    def solve(self, params: SolveParams) -> dict:
        st = SolveStats()
        self.call("fsl_pcg_solve", C.byref(params), C.byref(st))
        return st.as_dict()

    def pressure_update(self):
        self.call("fsl_pressure_update")

    def velocity_update(self):
        self.call("fsl_velocity_update")

    def extrapolate(self):
        self.call("fsl_extrapolate")

    def g2p(self):
        self.call("fsl_g2p")

    def step(self, dt, params: SolveParams) -> dict:
        st = SolveStats()
        self.call("fsl_step", float(dt), C.byref(params), C.byref(st))
        return st.as_dict()

    def advance_frame(self, cfl, frame_step, params: SolveParams) -> dict:
        """One frame of FlipDriver's loop with dt computed on the device."""
        fs = FrameStats()
        self.call("fsl_advance_frame", int(cfl), float(frame_step), C.byref(params), C.byref(fs))
        return fs.as_dict()

    def set_step_graph(self, enable=True):
        """advance_frame replays each step's front half from a HIP graph."""
        self.call("fsl_set_step_graph", int(enable))

    def project_forward(self, params: SolveParams) -> dict:
        st = SolveStats()
        self.call("fsl_project_forward", C.byref(params), C.byref(st))
        return st.as_dict()

    def project_backward(self, params: SolveParams, grad_out):
        """P^T grad_out. A float32 device tensor in box format is read in
        place and the result is a device tensor (no host staging); anything
        else goes through host memory as a numpy array."""
        if _is_device_tensor(grad_out):
            import torch

            g = grad_out.detach().to(torch.float32).contiguous()
            assert tuple(g.shape) == self.box + (3,)
            out = torch.empty_like(g)
            torch.cuda.current_stream(g.device).synchronize()
        else:
            g = np.ascontiguousarray(grad_out, dtype=np.float32)
            assert g.shape == self.box + (3,)
            out = np.empty_like(g)
        st = SolveStats()
        self.call("fsl_project_backward", C.byref(params), _ptr(g), _ptr(out), C.byref(st))
        return out, st.as_dict()

    def synchronize(self):
        self.call("fsl_synchronize")

    # -- rollouts (gradients with the recorded particle positions held fixed)
    def rollout_begin(self, max_substeps: int):
        self.call("fsl_rollout_begin", int(max_substeps))

    def _box3(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape == self.box + (3,)
        return a

    def rollout_step(self, dt, params: SolveParams, u0=None, want_u_post=False):
        u0 = self._box3(u0)
        up = np.empty(self.box + (3,), np.float32) if want_u_post else None
        st = SolveStats()
        self.call("fsl_rollout_step", float(dt), C.byref(params), _ptr(u0), _ptr(up),
                  C.byref(st))
        return (st.as_dict(), up) if want_u_post else st.as_dict()

    def rollout_replay(self, params: SolveParams, u0=None, force=None):
        u0 = self._box3(u0)
        f = np.ascontiguousarray(self.force if force is None else force, dtype=np.float32)
        up = np.empty(self.box + (3,), np.float32)
        st = SolveStats()
        self.call("fsl_rollout_replay", C.byref(params), _ptr(u0), _ptr(f), _ptr(up), C.byref(st))
        return up, st.as_dict()

    def rollout_backward(self, params: SolveParams, loss_grad):
        g = self._box3(loss_grad)
        gu = np.empty(self.box + (3,), np.float32)
        gf = np.zeros(3, np.float64)
        st = SolveStats()
        self.call("fsl_rollout_backward", C.byref(params), _ptr(g), _ptr(gu), _ptr(gf),
                  C.byref(st))
        return gu, gf, st.as_dict()

    def rollout_checkpoint(self, k: int, cap: int):
        pos = np.empty((cap, 3), np.float32)
        vel = np.empty((cap, 3), np.float32)
        perm = np.empty(cap, np.uint32)
        n, n_prev, dt = C.c_int64(), C.c_int64(), C.c_float()
        self.call("fsl_rollout_checkpoint", int(k), _ptr(pos), _ptr(vel), _ptr(perm),
                  C.byref(n), C.byref(n_prev), C.byref(dt))
        m = n.value
        return pos[:m].copy(), vel[:m].copy(), perm[:m].copy(), n_prev.value, dt.value

    # -- measurement hook
    def set_timing(self, enable=True):
        self.call("fsl_set_timing", int(enable))

    def timing(self) -> dict:
        out = (C.c_double * len(TIMING_NAMES))()
        self.call("fsl_get_timing", out, len(TIMING_NAMES))
        return dict(zip(TIMING_NAMES, list(out)))

    # -- PCG stage API
    def pcg_prepare(self, params: SolveParams):
        rr, rz = C.c_double(), C.c_double()
        self.call("fsl_pcg_prepare", C.byref(params), C.byref(rr), C.byref(rz))
        return rr.value, rz.value

    def pcg_iterate(self):
        v = [C.c_double() for _ in range(5)]
        self.call("fsl_pcg_iterate", *[C.byref(x) for x in v])
        return dict(zip(("pAp", "alpha", "rr", "rz", "beta"), (x.value for x in v)))

    def pcg_local1(self) -> float:
        """ComputeLocal1: Ap and the reduced pAp."""
        v = C.c_double()
        self.call("fsl_pcg_local1", C.byref(v))
        return v.value

    def pcg_local2(self):
        """ComputeLocal2: (alpha, rr, rz)."""
        v = [C.c_double() for _ in range(3)]
        self.call("fsl_pcg_local2", *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def pcg_local3(self) -> float:
        """ComputeLocal3: beta."""
        v = C.c_double()
        self.call("fsl_pcg_local3", C.byref(v))
        return v.value

    def pcg_finish(self) -> dict:
        st = SolveStats()
        self.call("fsl_pcg_finish", C.byref(st))
        return st.as_dict()

    # -- grids (box format over [-1, n])
    def get_vec(self, which=GRID_VELOCITY, device=None):
        """(values, activity) in box format: numpy arrays, or float32 / uint8
        tensors on `device` (a torch CUDA device of this context), written
        in place with no host staging."""
        if device is not None:
            import torch

            out = torch.empty(self.box + (3,), dtype=torch.float32, device=device)
            act = torch.empty(self.box, dtype=torch.uint8, device=device)
            torch.cuda.current_stream(out.device).synchronize()
        else:
            out = np.empty(self.box + (3,), np.float32)
            act = np.empty(self.box, np.uint8)
        self.call("fsl_get_vec", int(which), _ptr(out), _ptr(act))
        return out, act

    def set_vec(self, which, values, active):
        """Sets a vector grid from box-format values and activity: numpy
        arrays, or CUDA tensors of this context's device (read in place)."""
        if _is_device_tensor(values):
            import torch

            v = values.detach().to(torch.float32).contiguous()
            a = torch.as_tensor(active, device=v.device).to(torch.uint8).contiguous()
            assert tuple(v.shape) == self.box + (3,) and tuple(a.shape) == self.box
            torch.cuda.current_stream(v.device).synchronize()
        else:
            v = np.ascontiguousarray(values, dtype=np.float32)
            a = np.ascontiguousarray(active, dtype=np.uint8)
            assert v.shape == self.box + (3,) and a.shape == self.box
        self.call("fsl_set_vec", int(which), _ptr(v), _ptr(a))

    def get_scalar(self, which):
        out = np.empty(self.box, np.float32)
        self.call("fsl_get_scalar", int(which), _ptr(out))
        return out

    def get_scalar_noalloc(self, which):
        """get_scalar staged through the library's fixed device array (no
        allocation; boxes of at most 8,388,608 voxels)."""
        out = np.empty(self.box, np.float32)
        self.call("fsl_get_scalar_noalloc", int(which), _ptr(out))
        return out

    def set_scalar(self, which, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        self.call("fsl_set_scalar", int(which), _ptr(v))

    def get_int(self, which):
        out = np.empty(self.box, np.int32)
        self.call("fsl_get_int", int(which), _ptr(out))
        return out

    def set_marker(self, marker):
        m = np.ascontiguousarray(marker, dtype=np.int32)
        self.call("fsl_set_marker", _ptr(m))

    def active_leaves(self, fluid=False):
        n = C.c_int64()
        self.call("fsl_get_active_leaves", int(fluid), None, 0, C.byref(n))
        out = np.empty((n.value, 3), np.int32)
        self.call("fsl_get_active_leaves", int(fluid), _ptr(out), n.value, C.byref(n))
        return out

    def tile_rows(self):
        nt = int(np.prod(self.tiles))
        out = np.zeros(max(nt, 1), np.int64)
        self.call("fsl_tile_rows", out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out[:nt]
