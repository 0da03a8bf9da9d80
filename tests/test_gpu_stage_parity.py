"""PCG stage parity: fsl_pcg_prepare / fsl_pcg_local1 / local2 / local3
(PreparePCGSolver and ComputeLocal1-3, solver.cc:262-366) through the C ABI
against the oracle's FP32 solve in the reference's stage order
(solver_driver.cc:318-343), on the system of the 64^3 sphere drop's first
step (the oracle's grids after P2G, boundary conditions and gravity, set on
the GPU, so both sides assemble the same b).

  * With the dot products summed in the oracle's row order
    (FSL_DOT_ROW_ORDER), every scalar of every iteration (r.r and r.z after
    the preparation, then pAp, alpha, r.r, r.z, beta) and the final pressure
    are bit-identical to the oracle's, and so is the iteration count under
    the reference's stopping rule. This pins the SpMV, the Jacobi apply, the
    axpys and the scalar formulas to the oracle's arithmetic; only the
    summation order of the dots is a choice.
  * With the GPU's own tree-ordered FP32 dots, the iteration count is within
    +-2, and over the first 20 iterations each scalar sequence is as close to
    the FP64 one (the oracle's FP64 solve of the same system) as the
    oracle's sequential FP32 sequence is, within a factor 2: the two FP32
    orders round differently (measured up to 8e-4 apart in r.r by iteration
    13) and neither is the exact one.
  * Stage calls out of order return FSL_ERR_STATE.
"""
import numpy as np
import pytest

import fsl
from fsl import scenes
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

N = (64, 64, 64)


def first_step_system():
    """Marker and velocity of the sphere drop's first step just before
    ComputeDivergence, from the oracle, and the oracle itself."""
    mask = scenes.WORKLOADS["sphere_drop_64"][1](N)
    pos, vel = scenes.seed_particles(mask, 8, 1)
    o = Oracle(N)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    o.set_particles(pos, vel)
    dt, _ = o.compute_dt(1, 1.0 / 30, 0.0)
    o.move_particles(dt); o.clear(); o.p2g(); o.apply_bc(); o.save_velocities()
    o.delete_outside(); o.add_gravity(dt); o.apply_bc()
    return o


def gpu_with(o):
    ctx = fsl.Context(N, capacity=64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_marker(o.get_int(0))
    v, act = o.get_vec(0)
    ctx.set_vec(fsl.GRID_VELOCITY, v, act)
    ctx.divergence()
    o.divergence()
    assert np.array_equal(ctx.get_scalar(fsl.GRID_DIVERGENCE).view(np.uint32),
                          o.get_scalar(0).view(np.uint32))
    return ctx


def run_stages(ctx, prm, thr, max_it=1500):
    """The reference driver's loop over the stage calls; returns the
    per-iteration scalars and the iteration count."""
    rr, rz = ctx.pcg_prepare(prm)
    h = {"rr": [rr], "rz": [rz], "pAp": [], "alpha": [], "beta": []}
    it = 0
    while not (it >= max_it or np.float32(rr) < np.float32(thr)):
        pAp = ctx.pcg_local1()
        alpha, rr, rz = ctx.pcg_local2()
        beta = ctx.pcg_local3()
        for k, x in (("pAp", pAp), ("alpha", alpha), ("rr", rr), ("rz", rz), ("beta", beta)):
            h[k].append(x)
        it += 1
    st = ctx.pcg_finish()
    return {k: np.asarray(v, np.float64) for k, v in h.items()}, it, st


def test_stages_bit_exact_in_row_order():
    o = first_step_system()
    ctx = gpu_with(o)
    thr = float(np.float32(np.float32(np.prod(N)) * np.float32(0.5e-8)))
    rc, it_o, _ = o.solve(precision=0, stop_mode=0, threshold=thr)
    assert rc == 0
    ho = o.scalar_history()
    prm = fsl.solve_params(dims=N, variant=fsl.PCG_TWO_PASS, dot_order=fsl.DOT_ROW_ORDER)
    hg, it_g, st = run_stages(ctx, prm, thr)
    print(f"row order: {it_g} iterations (oracle {it_o}), {st['n_rows']} rows", flush=True)
    assert it_g == it_o
    for k in ("rr", "rz", "pAp", "alpha", "beta"):
        a = hg[k].astype(np.float32)
        b = ho[k].astype(np.float32)
        assert a.shape == b.shape, (k, a.shape, b.shape)
        bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad.size == 0, f"{k} differs first at iteration {bad[0]}: {a[bad[0]]} vs {b[bad[0]]}"
    p_g = ctx.get_scalar(fsl.GRID_PRESSURE)
    p_o = o.get_scalar(1)
    assert np.array_equal(p_g.view(np.uint32), p_o.view(np.uint32))
    # The whole-solve entry point takes the same path.
    st2 = ctx.solve(prm)
    assert st2["iterations"] == it_o
    assert np.array_equal(ctx.get_scalar(fsl.GRID_PRESSURE).view(np.uint32), p_o.view(np.uint32))
    ctx.close()


def test_stages_tree_order_first_20_iterations():
    o = first_step_system()
    ctx = gpu_with(o)
    thr = float(np.float32(np.float32(np.prod(N)) * np.float32(0.5e-8)))
    rc, it_o, _ = o.solve(precision=0, stop_mode=0, threshold=thr)
    assert rc == 0
    ho = o.scalar_history()
    prm = fsl.solve_params(dims=N, variant=fsl.PCG_TWO_PASS)
    hg, it_g, _ = run_stages(ctx, prm, thr)
    print(f"tree order: {it_g} iterations (oracle {it_o})", flush=True)
    assert abs(it_g - it_o) <= 2
    # The FP64 sequence of the same system (the oracle, same stage order).
    o.set_scalar(1, np.zeros(o.box, np.float32))
    rc, _, _ = o.solve(precision=1, stop_mode=0, threshold=thr)
    assert rc == 0
    h64 = o.scalar_history()
    report = {}
    for k in ("rr", "rz", "pAp", "alpha", "beta"):
        m = 20
        ref = h64[k][:m]
        e_gpu = np.linalg.norm(hg[k][:m] - ref) / np.linalg.norm(ref)
        e_orc = np.linalg.norm(ho[k][:m] - ref) / np.linalg.norm(ref)
        report[k] = (float(e_gpu), float(e_orc))
        assert e_gpu <= 2.0 * e_orc + 1e-6, (k, e_gpu, e_orc)
    print("rel. L2 distance to FP64 over the first 20 iterations (gpu tree, oracle):", report)
    ctx.close()


def test_stage_calls_out_of_order():
    o = first_step_system()
    ctx = gpu_with(o)
    prm = fsl.solve_params(dims=N, variant=fsl.PCG_TWO_PASS)
    ctx.pcg_prepare(prm)
    with pytest.raises(fsl.FslError) as e:
        ctx.pcg_local2()
    assert e.value.code == -3
    ctx.pcg_local1()
    with pytest.raises(fsl.FslError) as e:
        ctx.pcg_local3()
    assert e.value.code == -3
    ctx.pcg_local2()
    ctx.pcg_local3()
    # Row order is a two-pass feature of the FP32 and FP64 solves.
    bad = fsl.solve_params(precision=fsl.PCG_MIXED, dims=N, dot_order=fsl.DOT_ROW_ORDER)
    with pytest.raises(fsl.FslError) as e:
        ctx.solve(bad)
    assert e.value.code == -1
    ctx.close()
