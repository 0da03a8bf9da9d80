"""The 64^3 sphere drop for its full 20 frames (pool y < 16, sphere r = 10
at (32, 46, 32), 8 particles per cell, gravity 9.8, flip 0.95, cfl 1,
30 frames/s; the frame loop of FlipDriver, flip_app_sim.cc:357-428), GPU
against the oracle with the reference's FP32 solve and stopping rule.

Each frame starts both sides from the oracle's particles (the reference's
trajectory; the GPU's particles would otherwise drift from it by rounding
and cells would eventually differ at the last bit of a position), runs the
frame's steps with dt from each side's own CFL, and checks:
  * the frame's dt sequence: bit-equal;
  * after every step, labels and per-cell particle counts: bit-exact, and
    the particles in the same order with bit-equal positions;
  * PCG iterations of every step: within +-2 of the oracle's.
A free-running GPU trajectory over the same 20 frames keeps every particle
and ends with the same fluid-cell count as the oracle's to 1%.
"""
import numpy as np
import pytest

import fsl
from fsl import scenes
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

N = (64, 64, 64)
FRAMES = 20
FRAME_STEP = np.float32(1.0 / 30)


def scene():
    mask = scenes.WORKLOADS["sphere_drop_64"][1](N)
    return scenes.seed_particles(mask, 8, 1)


def new_pair(pos, vel):
    ctx = fsl.Context(N, capacity=pos.shape[0] + 64)
    orc = Oracle(N)
    for s in (ctx, orc):
        s.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
        s.set_box_boundary()
        s.set_particles(pos, vel)
    return ctx, orc


def test_sphere_drop_20_frames_lockstep():
    pos, vel = scene()
    ctx, orc = new_pair(pos, vel)
    prm = fsl.solve_params(dims=N)
    total_steps, worst = 0, 0
    for f in range(FRAMES):
        op, ovel, ogv, _ = orc.get_particles()
        ctx.set_particles(op, ovel, ogv)
        step_o = step_g = np.float32(0.0)
        while True:
            dt, step_o = orc.compute_dt(1, FRAME_STEP, step_o)
            gdt, step_g = ctx.compute_dt(1, FRAME_STEP, step_g)
            assert gdt == dt and step_g == step_o, (f, gdt, dt)
            st = ctx.step(dt, prm)
            rc, it = orc.run_one_step(dt)
            assert rc == 0
            assert abs(st["iterations"] - it) <= 2, (f, st["iterations"], it)
            worst = max(worst, abs(st["iterations"] - it))
            assert np.array_equal(ctx.get_int(fsl.GRID_MARKER), orc.get_int(0)), f
            assert np.array_equal(ctx.get_int(fsl.GRID_CELL_COUNT), orc.cell_counts()), f
            gp, _, _ = ctx.get_particles()
            op, _, _, _ = orc.get_particles()
            assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f
            total_steps += 1
            if step_o + np.float32(1e-5) >= FRAME_STEP:
                break
    print(f"{FRAMES} frames, {total_steps} steps, iteration counts within {worst}")
    assert total_steps >= FRAMES
    ctx.close()


def test_sphere_drop_20_frames_free_running():
    pos, vel = scene()
    ctx, orc = new_pair(pos, vel)
    prm = fsl.solve_params(dims=N)
    for side in (ctx, orc):
        for f in range(FRAMES):
            step = np.float32(0.0)
            while True:
                dt, step = side.compute_dt(1, FRAME_STEP, step)
                if side is ctx:
                    side.step(dt, prm)
                else:
                    assert side.run_one_step(dt)[0] == 0
                if step + np.float32(1e-5) >= FRAME_STEP:
                    break
    assert ctx.num_particles() == orc.num_particles() == pos.shape[0]
    fg = int((ctx.get_int(fsl.GRID_MARKER) == 1).sum())
    fo = int((orc.get_int(0) == 1).sum())
    print("fluid cells after 20 frames: gpu", fg, "oracle", fo)
    assert abs(fg - fo) <= 0.01 * fo
    ctx.close()
