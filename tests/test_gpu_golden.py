"""The library against the committed fixtures of tests/golden/make_golden.py
(answers computed there with numpy and scipy, independently of the oracle):
the FP32 divergence bit for bit, and the pressure of FP64 solves (Jacobi
single-reduction and two-pass, block IC(0)) against the dense FP64 solution,
to the FP32 rounding of the stored pressure. Mixed and FP32-IC solves reach
their stated true residual."""
import os

import numpy as np
import pytest

import fsl

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "projection_systems.npz")


def golden_system(s):
    g = np.load(GOLDEN)
    return {k: g[f"{k}_{s}"] for k in ("dims", "tiles", "marker", "faces", "b", "x")}


@pytest.mark.parametrize("s", range(3))
def test_gpu_matches_golden_systems(s):
    g = golden_system(s)
    n = tuple(int(d) for d in g["dims"])
    tiles = tuple(int(t) for t in g["tiles"])
    print(f"golden system {s}: dims {n}, tiles {tiles}, {g['b'].size} rows", flush=True)
    rows = g["marker"] == 1
    ctx = fsl.Context(n, tiles=tiles, capacity=64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_marker(g["marker"].astype(np.int32))
    ctx.set_vec(fsl.GRID_VELOCITY, g["faces"], np.ones(g["marker"].shape, np.uint8))
    ctx.apply_bc()
    ctx.divergence()
    assert np.array_equal(ctx.get_scalar(fsl.GRID_DIVERGENCE)[rows], g["b"])
    x = g["x"]
    tight = dict(stop_mode=fsl.STOP_REL, threshold=1e-12, max_iterations=5000, dims=n)
    for name, prm in {
        "fp64 single-reduction": fsl.solve_params(precision=fsl.PCG_FP64, **tight),
        "fp64 two-pass": fsl.solve_params(precision=fsl.PCG_FP64, variant=fsl.PCG_TWO_PASS,
                                          **tight),
        "fp64 block IC(0)": fsl.solve_params(precision=fsl.PCG_FP64,
                                             preconditioner=fsl.PRECOND_IC, **tight),
    }.items():
        st = ctx.solve(prm)
        p = ctx.get_scalar(fsl.GRID_PRESSURE)[rows].astype(np.float64)
        err = np.linalg.norm(p - x) / np.linalg.norm(x)
        print(f"  {name}: {st['iterations']} iterations, pressure vs dense {err:.2e}")
        assert st["true_rel_residual"] <= 2e-12 and err <= 1e-6
    loose = dict(stop_mode=fsl.STOP_REL, threshold=1e-6, max_iterations=5000, dims=n)
    for name, prm in {
        "mixed": fsl.solve_params(precision=fsl.PCG_MIXED, **loose),
        "mixed block IC(0)": fsl.solve_params(precision=fsl.PCG_MIXED,
                                              preconditioner=fsl.PRECOND_IC, **loose),
    }.items():
        st = ctx.solve(prm)
        print(f"  {name}: {st['iterations']} iterations, true rel residual "
              f"{st['true_rel_residual']:.2e}")
        assert st["true_rel_residual"] <= 1e-6
    ctx.close()
