"""Block IC(0) preconditioner (the IC branch of BuildPreconditioner,
solver.cc:103-106, with openvdb.patch:209-224) on the GPU against the oracle's
restatement of OpenVDB 4.0.1's IncompleteCholeskyPreconditioner.

Scenes: 32^3, a fluid block y < 20 with 6% random SOLID and 4% AIR cells and
i.i.d. N(0,1) face velocities; every FLUID component touches AIR. Tiles
(1,1,1), (2,2,2) and (1,2,4): the factor is block-diagonal over tiles.

  * z = M^-1 b: r.z after preparation (x0 = 0, so r = b) equals b . z_oracle,
    and one PCG iteration gives x = alpha z, equal to the oracle's to 1e-5
    (the FP32 sums of pAp differ in order only);
  * full solves: FP32 with the reference's rule and FP64 to 1e-10 take the
    oracle's iteration count (+-2) and match its pressure (1e-5 / 1e-6); MIXED
    reaches a true relative residual <= 1e-6; IC needs fewer iterations than
    Jacobi;
  * a sealed two-cell FLUID pocket (singular block) fails the factor's pivot
    test: FSL_ERR_PRECONDITIONER, as the oracle's build fails.
"""
import numpy as np
import pytest
from scipy import ndimage

import fsl
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

N = (32, 32, 32)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def scene(n, seed):
    rng = np.random.default_rng(seed)
    box = tuple(d + 2 for d in n)
    inner = np.zeros(n, np.int32)
    inner[:, : n[1] * 5 // 8, :] = 1
    u = rng.random(n)
    inner[u < 0.06] = 2
    inner[(u >= 0.06) & (u < 0.10)] = 0
    marker = np.full(box, 2, np.int32)
    marker[1:-1, 1:-1, 1:-1] = inner
    lab, k = ndimage.label(marker == 1)
    air = ndimage.binary_dilation(marker == 0)
    for comp in range(1, k + 1):
        if not np.any(air[lab == comp]):
            marker[lab == comp] = 0
    faces = rng.standard_normal(box + (3,)).astype(np.float32)
    return marker, faces


def oracle_for(n, marker, faces, fp64=False):
    o = Oracle(n, fp64=fp64)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    o.set_marker(marker)
    o.set_vec(0, faces, np.ones(marker.shape, np.uint8))
    o.apply_bc()
    o.divergence()
    return o


def ctx_for(n, tiles, marker, faces):
    ctx = fsl.Context(n, tiles=tiles, capacity=64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_marker(marker)
    ctx.set_vec(fsl.GRID_VELOCITY, faces, np.ones(marker.shape, np.uint8))
    ctx.apply_bc()
    ctx.divergence()
    return ctx


def ic_params(**kw):
    kw.setdefault("dims", N)
    return fsl.solve_params(preconditioner=fsl.PRECOND_IC, **kw)


@pytest.mark.parametrize("tiles", [(1, 1, 1), (2, 2, 2), (1, 2, 4)])
def test_ic_apply_matches_oracle(tiles):
    print(f"ic apply, tiles {tiles}", flush=True)
    marker, faces = scene(N, 31)
    rows = marker == 1
    o = oracle_for(N, marker, faces)
    b = o.get_scalar(0)[rows].astype(np.float64)
    d, z = o.ic_factor(tiles, b)
    assert d is not None and len(z) == rows.sum()
    ctx = ctx_for(N, tiles, marker, faces)
    assert np.array_equal(ctx.get_scalar(fsl.GRID_DIVERGENCE)[rows], o.get_scalar(0)[rows])
    rr, rz = ctx.pcg_prepare(ic_params(threshold=0.0))
    assert abs(rr - float(b @ b)) <= 1e-6 * float(b @ b)
    assert abs(rz - float(b @ z)) <= 1e-5 * abs(float(b @ z))
    # One iteration: x = alpha z0 with z0 = M^-1 b.
    ctx.solve(ic_params(threshold=0.0, max_iterations=1))
    p = ctx.get_scalar(fsl.GRID_PRESSURE)[rows].astype(np.float64)
    rc, it, _ = o.solve(precision=0, stop_mode=0, threshold=0.0, max_iter=1, precond=1,
                        tiles=tiles)
    assert rc == 0 and it == 1
    p_or = o.get_scalar(1)[rows].astype(np.float64)
    err = rel(p, p_or)
    c = float(p @ z) / float(z @ z)
    print(f"  one iteration: pressure vs oracle {err:.2e}, vs alpha z {rel(p, c * z):.2e}")
    assert err <= 1e-5 and rel(p, c * z) <= 1e-5


@pytest.mark.parametrize("tiles", [(1, 1, 1), (2, 2, 2)])
def test_ic_solves_match_oracle(tiles):
    print(f"ic solves, tiles {tiles}", flush=True)
    marker, faces = scene(N, 33)
    rows = marker == 1
    ctx = ctx_for(N, tiles, marker, faces)

    # FP32, the reference's rule r.r < N^3 * 0.5e-8.
    o = oracle_for(N, marker, faces)
    rc, it_or, _ = o.solve(precision=0, precond=1, tiles=tiles)
    assert rc == 0
    st = ctx.solve(ic_params(precision=fsl.PCG_FP32))
    p = ctx.get_scalar(fsl.GRID_PRESSURE)[rows]
    err = rel(p, o.get_scalar(1)[rows])
    print(f"  fp32: {st['iterations']} iterations (oracle {it_or}), pressure {err:.2e}")
    assert abs(st["iterations"] - it_or) <= 2 and err <= 1e-5

    # FP64 to a relative residual of 1e-10, against the FP64 oracle.
    o64 = oracle_for(N, marker, faces, fp64=True)
    rc, it64, _ = o64.solve(precision=1, stop_mode=1, threshold=1e-10, max_iter=5000,
                            precond=1, tiles=tiles)
    assert rc == 0
    st = ctx.solve(ic_params(precision=fsl.PCG_FP64, stop_mode=fsl.STOP_REL, threshold=1e-10,
                             max_iterations=5000))
    p64 = ctx.get_scalar(fsl.GRID_PRESSURE)[rows]
    err = rel(p64, o64.get_scalar(1)[rows])
    print(f"  fp64: {st['iterations']} iterations (oracle {it64}), pressure {err:.2e}")
    assert abs(st["iterations"] - it64) <= 2 and err <= 1e-6

    # Jacobi needs more iterations on the same system.
    st_j = ctx.solve(fsl.solve_params(precision=fsl.PCG_FP64, stop_mode=fsl.STOP_REL,
                                      threshold=1e-10, max_iterations=5000,
                                      variant=fsl.PCG_TWO_PASS))
    print(f"  jacobi fp64: {st_j['iterations']} iterations")
    assert st_j["iterations"] > st["iterations"]

    # MIXED (two-pass with the FP64 refresh) to 1e-6.
    st = ctx.solve(ic_params(precision=fsl.PCG_MIXED, stop_mode=fsl.STOP_REL, threshold=1e-6,
                             max_iterations=5000))
    pm = ctx.get_scalar(fsl.GRID_PRESSURE)[rows]
    print(f"  mixed: {st['iterations']} iterations, true rel residual "
          f"{st['true_rel_residual']:.2e}, pressure vs fp64 {rel(pm, p64):.2e}")
    assert st["true_rel_residual"] <= 1e-6 and rel(pm, p64) <= 1e-5

    # MIXED with a refresh every 4 iterations: each refresh recomputes the
    # residual in FP64 and applies IC to it in place of B's apply.
    st = ctx.solve(ic_params(precision=fsl.PCG_MIXED, stop_mode=fsl.STOP_REL, threshold=1e-6,
                             max_iterations=5000, refresh_every=4))
    pm4 = ctx.get_scalar(fsl.GRID_PRESSURE)[rows]
    print(f"  mixed, refresh every 4: {st['iterations']} iterations, true rel residual "
          f"{st['true_rel_residual']:.2e}, pressure vs fp64 {rel(pm4, p64):.2e}")
    assert st["iterations"] > 8
    assert st["true_rel_residual"] <= 1e-6 and rel(pm4, p64) <= 1e-5


def test_ic_stage_api_matches_solve():
    print("ic stage api", flush=True)
    marker, faces = scene(N, 35)
    ctx = ctx_for(N, (1, 1, 1), marker, faces)
    prm = ic_params(precision=fsl.PCG_FP32)
    st = ctx.solve(prm)
    p = ctx.get_scalar(fsl.GRID_PRESSURE)
    # The stage API runs the same kernels one iteration at a time.
    ctx.pcg_prepare(prm)
    for _ in range(st["iterations"]):
        ctx.pcg_iterate()
    s = ctx.pcg_finish()
    assert s["iterations"] == st["iterations"]
    assert np.array_equal(ctx.get_scalar(fsl.GRID_PRESSURE), p)


def test_ic_sealed_pocket_fails_the_pivot_test():
    print("ic sealed pocket", flush=True)
    marker, faces = scene(N, 37)
    # Two FLUID cells walled in by SOLID in the air region: a singular block.
    marker[1:-1, 24:30, 1:-1] = 0
    marker[19:24, 26:29, 19:22] = 2
    marker[20, 27, 20] = marker[21, 27, 20] = 1
    o = oracle_for(N, marker, faces)
    rc, _, _ = o.solve(precision=0, precond=1)
    assert rc == -7
    ctx = ctx_for(N, (1, 1, 1), marker, faces)
    with pytest.raises(fsl.FslError) as e:
        ctx.solve(ic_params(precision=fsl.PCG_FP32))
    assert e.value.code == -15
    with pytest.raises(fsl.FslError) as e:
        ctx.solve(ic_params(precision=fsl.PCG_FP32, variant=fsl.PCG_SINGLE_REDUCTION))
    assert e.value.code == -1  # FSL_ERR_ARG: IC runs with the two-pass variant only
