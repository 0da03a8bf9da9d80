"""The C++ host classes (include/fsl.hpp) driving libfsl: fsl_water_app runs
one frame of the 32^3 sphere drop through FlipSimulation / PCGDriver /
FlipDriver and writes the reference's binary particle frame
(particles.cc:166-200); the same frame run through the Python binding must
give the same particles bit for bit (same library, deterministic kernels)."""
import os
import subprocess

import numpy as np
import pytest

import fsl
from fsl import scenes

pytestmark = pytest.mark.gpu

APP = os.path.join(os.path.dirname(fsl.__file__), os.pardir, "fsl_water_app")


def read_frame(path):
    raw = open(path, "rb").read()
    head = np.frombuffer(raw[:28], dtype=np.int32)
    start, end, count = head[:3], head[3:6], int(head[6])
    rec = np.frombuffer(raw[28:], dtype=np.dtype([("id", "<i4"), ("pos", "<f4", 3),
                                                   ("vel", "<f4", 3)]))
    assert rec.shape[0] == count
    return start, end, rec


def test_water_app_matches_python_path(tmp_path):
    n = 32
    out = str(tmp_path / "particles_")
    r = subprocess.run([APP, f"n={n}", "scene=sphere_drop", "frames=1", "seed=1",
                        f"output={out}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    start, end, rec = read_frame(out + "00000001")
    assert list(start) == [0, 0, 0] and list(end) == [n - 1] * 3
    assert np.array_equal(rec["id"], np.arange(rec.shape[0]))

    dims = (n, n, n)
    mask = scenes.sphere_drop_mask(dims, 8, (16, 23, 16), 5)
    ctx = fsl.Context(dims, capacity=int(mask.sum()) * 8 + 64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.seed_particles(mask, 8, 1)
    prm = fsl.solve_params(dims=dims)
    frame = np.float32(1.0 / 30)
    step = np.float32(0.0)
    steps = 0
    while np.float32(step + np.float32(1e-5)) < frame:
        dt, step = ctx.compute_dt(1, frame, step)
        step = np.float32(step)
        ctx.step(dt, prm)
        steps += 1
    pos, vel, _ = ctx.get_particles()
    ctx.close()
    assert steps >= 1
    assert pos.shape[0] == rec.shape[0]
    assert np.array_equal(pos, rec["pos"]) and np.array_equal(vel, rec["vel"])
