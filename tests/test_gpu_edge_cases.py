"""Edge cases of the step and the solver through the C ABI: an empty domain,
every particle leaving an open tank, and argument errors (returned as
FSL_ERR_* codes, never aborts, as the drop-in boundary requires)."""
import numpy as np
import pytest

import fsl
from fsl import scenes
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

PARAMS = {
    "fp32_reference_rule": dict(precision=fsl.PCG_FP32),
    "mixed": dict(precision=fsl.PCG_MIXED, stop_mode=fsl.STOP_REL, threshold=1e-6),
    "fp64_two_pass": dict(precision=fsl.PCG_FP64, stop_mode=fsl.STOP_REL, threshold=1e-6,
                          variant=fsl.PCG_TWO_PASS),
    "ic": dict(precision=fsl.PCG_FP32, preconditioner=fsl.PRECOND_IC),
}


@pytest.mark.parametrize("name", list(PARAMS))
def test_empty_domain_steps(name):
    print(f"empty domain, {name}", flush=True)
    n = (16, 24, 16)
    ctx = fsl.Context(n, capacity=64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    prm = fsl.solve_params(dims=n, **PARAMS[name])
    for _ in range(2):
        st = ctx.step(1.0 / 30, prm)
        assert st["iterations"] == 0 and st["n_rows"] == 0
    assert ctx.num_particles() == 0
    assert not np.any(ctx.get_scalar(fsl.GRID_PRESSURE))
    assert not np.any(ctx.get_int(fsl.GRID_MARKER) == 1)
    ctx.close()
    o = Oracle(n)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    rc, it = o.run_one_step(1.0 / 30)
    assert rc == 0 and it == 0


def test_all_particles_leave_the_tank():
    print("all particles leave", flush=True)
    n = (16, 16, 16)
    mask = np.zeros(n, np.uint8)
    mask[4:12, 14:16, 4:12] = 1  # the top two layers
    pos, vel = scenes.seed_particles(mask, 4, 5)
    vel[:, 1] = 200.0  # 20 cells up in one step of 0.1
    ctx = fsl.Context(n, capacity=pos.shape[0])
    o = Oracle(n)
    for s in (ctx, o):
        s.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
        s.set_box_boundary(True)  # open top
        s.set_particles(pos, vel)
    st = ctx.step(0.1, fsl.solve_params(dims=n))
    rc, it = o.run_one_step(0.1)
    assert rc == 0
    assert ctx.num_particles() == o.num_particles() == 0
    assert st["iterations"] == it == 0
    assert np.array_equal(ctx.get_int(fsl.GRID_MARKER), o.get_int(0))
    ctx.close()


def test_argument_errors():
    print("argument errors", flush=True)
    for dims, tiles in (((12, 16, 16), (1, 1, 1)), ((16, 16, 16), (3, 1, 1)),
                        ((16, 16, 16), (0, 1, 1)), ((16, 16, 8), (1, 1, 2))):
        with pytest.raises(fsl.FslError) as e:
            fsl.Context(dims, tiles=tiles)
        assert e.value.code == -1
    with pytest.raises(fsl.FslError) as e:
        fsl.Context((16, 16, 16), capacity=1 << 32)  # more particles than 32-bit slots
    assert e.value.code == -1
    n = (16, 16, 16)
    ctx = fsl.Context(n, capacity=64)
    ctx.set_box_boundary()
    for bad in (dict(precision=3), dict(stop_mode=2), dict(variant=3), dict(preconditioner=2),
                dict(preconditioner=fsl.PRECOND_IC, variant=fsl.PCG_SINGLE_REDUCTION)):
        with pytest.raises(fsl.FslError) as e:
            ctx.solve(fsl.solve_params(dims=n, **bad))
        assert e.value.code == -1, bad
    # The context still works after the rejected calls.
    st = ctx.step(1.0 / 30, fsl.solve_params(dims=n))
    assert st["iterations"] == 0
    ctx.close()
