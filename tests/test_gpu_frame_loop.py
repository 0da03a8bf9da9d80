"""The device-resident frame loop (fsl_advance_frame: FlipDriver's While
loop, flip_app_sim.cc:357-462, with dt computed on the device) against the
same loop driven from the host (fsl_compute_dt, a host read of the CFL
maximum, then fsl_step): the same dt sequence, step counts and bit-identical
particles and pressure. A fast scene (initial velocity 40 voxel/s, so
v_inf > 30 and a frame takes two steps or more) exercises the dt clamps:
cfl / v_inf, the cut at the frame's end and the snap past 90% of it. With
the step graph on, the front half of every step after the first is replayed
from a captured HIP graph (both particle buffers, so both graphs), with the
same results.
"""
import numpy as np
import pytest

import fsl
from fsl import scenes

pytestmark = pytest.mark.gpu

N = (32, 32, 32)
FRAME = np.float32(1.0 / 30)


def make(vel0):
    mask = scenes.sphere_drop_mask(N, 8, (16, 22, 16), 6)
    pos, vel = scenes.seed_particles(mask, 8, 21)
    vel[:] = vel0
    ctx = fsl.Context(N, capacity=pos.shape[0] + 64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_particles(pos, vel)
    return ctx


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("vel0", [(0.0, 0.0, 0.0), (40.0, -25.0, 10.0)])
def test_device_frame_loop_matches_host_loop(vel0, graph):
    prm = fsl.solve_params(precision=fsl.PCG_MIXED, stop_mode=fsl.STOP_REL, threshold=1e-6,
                           max_iterations=5000)
    a, b = make(vel0), make(vel0)
    a.set_step_graph(graph)
    steps_a = steps_b = 0
    for f in range(3):
        fs = a.advance_frame(1, FRAME, prm)
        steps_a += fs["steps"]
        step = np.float32(0.0)
        n = 0
        while True:
            dt, step = b.compute_dt(1, FRAME, step)
            b.step(dt, prm)
            n += 1
            if np.float64(step) + 1e-5 >= np.float64(FRAME):
                break
        steps_b += n
        assert fs["steps"] == n, (f, fs["steps"], n)
        assert fs["dt_last"] == dt
        print(f"frame {f}: {n} steps, last dt {dt}")
    assert steps_a == steps_b
    if vel0[0]:
        assert steps_a > 3
    pa, va, ga = a.get_particles()
    pb, vb, gb = b.get_particles()
    for x, y in ((pa, pb), (va, vb), (ga, gb)):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.array_equal(a.get_scalar(fsl.GRID_PRESSURE).view(np.uint32),
                          b.get_scalar(fsl.GRID_PRESSURE).view(np.uint32))
    assert a.num_particles() == b.num_particles()
    a.close(); b.close()
