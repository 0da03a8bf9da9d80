"""Predicted single-reduction solves (fsl_set_step_graph on: after a first
solve, a solve enqueues as many iterations as the previous one took, then
its convergence check, and reads the state once) against the same solves
with the solver state read after every batch: the same iteration counts and
bit-identical pressures, when the prediction is too low (the solve continues
32 iterations at a time), too high (the launches past convergence exit) and
right, for the mixed and the FP64 solve. The system is the 48x40x32 sphere
drop's first step.
"""
import numpy as np
import pytest

import fsl
from fsl import scenes

pytestmark = pytest.mark.gpu

N = (48, 40, 32)


def ready_context(graph):
    mask = scenes.sphere_drop_mask(N, 10, (24, 27, 16), 7)
    pos, vel = scenes.seed_particles(mask, 8, 5)
    ctx = fsl.Context(N, capacity=pos.shape[0] + 64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_particles(pos, vel)
    dt, _ = ctx.compute_dt(1, np.float32(1.0 / 30), 0.0)
    ctx.move_particles(dt); ctx.clear(); ctx.p2g(); ctx.apply_bc(); ctx.save_velocities()
    ctx.delete_outside(); ctx.add_gravity(dt); ctx.apply_bc(); ctx.divergence()
    ctx.set_step_graph(graph)
    return ctx


@pytest.mark.parametrize("precision", [fsl.PCG_MIXED, fsl.PCG_FP64])
def test_predicted_solves_match_batch_checked_solves(precision):
    # (first threshold, second threshold): the first solve sets the
    # prediction the second one replays.
    cases = [(1e-3, 1e-9), (1e-9, 1e-4), (1e-6, 1e-6)]
    a, b = ready_context(True), ready_context(False)
    for t1, t2 in cases:
        for thr in (t1, t2):
            prm = fsl.solve_params(precision=precision, stop_mode=fsl.STOP_REL, threshold=thr,
                                   max_iterations=5000, variant=fsl.PCG_SINGLE_REDUCTION)
            sa = a.solve(prm)
            sb = b.solve(prm)
            print(f"threshold {thr:g}: predicted {sa['iterations']} iterations, "
                  f"checked per batch {sb['iterations']}; true rel. residual "
                  f"{sa['true_rel_residual']:.3e}")
            assert sa["iterations"] == sb["iterations"], (thr, sa["iterations"], sb["iterations"])
            # The mixed solve confirms on the true residual; the FP64 one stops
            # on the recursive residual.
            assert sa["true_rel_residual"] <= thr * (1.0001 if precision == fsl.PCG_MIXED else 2.0)
            pa = a.get_scalar(fsl.GRID_PRESSURE)
            pb = b.get_scalar(fsl.GRID_PRESSURE)
            assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), thr
    a.close(); b.close()
