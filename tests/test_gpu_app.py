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


@pytest.mark.parametrize("pc", ["jacobi", "ic"])
def test_water_app_matches_python_path(tmp_path, pc):
    n = 32
    out = str(tmp_path / "particles_")
    r = subprocess.run([APP, f"n={n}", "scene=sphere_drop", "frames=1", "seed=1",
                        f"output={out}", f"preconditioner={pc}"], capture_output=True, text=True,
                       timeout=120)
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
    prm = fsl.solve_params(dims=dims, preconditioner=fsl.PRECOND_IC if pc == "ic"
                           else fsl.PRECOND_JACOBI)
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


def test_water_app_device_loop_matches_host_loop(tmp_path):
    """FlipDriver::NextFrameOnDevice (the frame on the device, step graph on)
    writes the same particle frames as FlipDriver::NextFrame, bit for bit,
    over three frames of the 32^3 sphere drop."""
    frames = {}
    for dl in ("0", "1"):
        out = str(tmp_path / f"p{dl}_")
        r = subprocess.run([APP, "n=32", "scene=sphere_drop", "frames=3", "seed=2",
                            f"output={out}", f"device_loop={dl}"], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        print(r.stdout.strip())
        frames[dl] = [read_frame(out + f"{f:08d}")[2] for f in (1, 2, 3)]
    for a, b in zip(frames["0"], frames["1"]):
        assert a.shape == b.shape
        assert np.array_equal(a["pos"].view(np.uint32), b["pos"].view(np.uint32))
        assert np.array_equal(a["vel"].view(np.uint32), b["vel"].view(np.uint32))


def vdb_key(i, j, k):
    """OpenVDB Tree3<T,3,3> value-iteration order (root by origin, then leaf,
    then voxel offsets, each x-major)."""
    i, j, k = (np.asarray(a, np.int64) for a in (i, j, k))
    r = (((i >> 6) + 1024) * 2048 + ((j >> 6) + 1024)) * 2048 + ((k >> 6) + 1024)
    leaf = (((i >> 3) & 7) << 6) | (((j >> 3) & 7) << 3) | ((k >> 3) & 7)
    vox = ((i & 7) << 6) | ((j & 7) << 3) | (k & 7)
    return (r << 18) | (leaf << 9) | vox


def read_grid(path, ncomp):
    lines = open(path).read().splitlines()
    start = [int(x) for x in lines[0].split()]
    end = [int(x) for x in lines[1].split()]
    ijk, vals = [], []
    for ln in lines[2:]:
        a, b = ln.split(":")
        ijk.append([int(x) for x in a.split()])
        vals.append([float(x) for x in b.split()])
    return start, end, np.array(ijk, np.int64).reshape(-1, 3), np.array(vals).reshape(-1, ncomp)


def test_water_app_grid_dumps(tmp_path):
    """Text grid dumps (vector_grid.cc:133-160, scalar_grid.cc:134-160): active
    voxels in OpenVDB iteration order, values printed with %f."""
    n = 32
    gout = str(tmp_path / "grid_")
    r = subprocess.run([APP, f"n={n}", "scene=sphere_drop", "frames=1", "seed=1",
                        f"grid_output={gout}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dims = (n, n, n)
    mask = scenes.sphere_drop_mask(dims, 8, (16, 23, 16), 5)
    ctx = fsl.Context(dims, capacity=int(mask.sum()) * 8 + 64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.seed_particles(mask, 8, 1)
    prm = fsl.solve_params(dims=dims)
    frame = np.float32(1.0 / 30)
    step = np.float32(0.0)
    while np.float32(step + np.float32(1e-5)) < frame:
        dt, step = ctx.compute_dt(1, frame, step)
        step = np.float32(step)
        ctx.step(dt, prm)
    v, act = ctx.get_vec(fsl.GRID_VELOCITY)
    marker = ctx.get_int(fsl.GRID_MARKER)
    pres = ctx.get_scalar(fsl.GRID_PRESSURE)
    ctx.close()
    # velocity: active faces over [0, n] per axis
    start, end, ijk, vals = read_grid(gout + "velocity_00000001", 3)
    assert start == [0, 0, 0] and end == [n - 1] * 3
    sub = act[1:n + 2, 1:n + 2, 1:n + 2]
    want = np.argwhere(sub)
    order = np.argsort(vdb_key(want[:, 0], want[:, 1], want[:, 2]), kind="stable")
    want = want[order]
    assert np.array_equal(ijk, want)
    ref = v[want[:, 0] + 1, want[:, 1] + 1, want[:, 2] + 1]
    assert np.allclose(vals, ref.astype(np.float64), rtol=0, atol=5e-7)  # printed with %f
    # pressure: FLUID cells
    _, _, ijk, vals = read_grid(gout + "pressure_00000001", 1)
    fl = np.argwhere(marker[1:n + 1, 1:n + 1, 1:n + 1] == 1)
    fl = fl[np.argsort(vdb_key(fl[:, 0], fl[:, 1], fl[:, 2]), kind="stable")]
    assert np.array_equal(ijk, fl)
    assert np.allclose(vals[:, 0], pres[fl[:, 0] + 1, fl[:, 1] + 1, fl[:, 2] + 1], atol=5e-7)
    # marker: FLUID or SOLID cells
    _, _, ijk, vals = read_grid(gout + "marker_00000001", 1)
    ms = marker[1:n + 1, 1:n + 1, 1:n + 1]
    on = np.argwhere((ms == 1) | (ms == 2))
    on = on[np.argsort(vdb_key(on[:, 0], on[:, 1], on[:, 2]), kind="stable")]
    assert np.array_equal(ijk, on)
    assert np.array_equal(vals[:, 0], ms[on[:, 0], on[:, 1], on[:, 2]].astype(np.float64))


def test_water_app_save_data_to_file(tmp_path):
    """SaveDataToFile / NextFrame (flip_simulation.cc:1284-1322) as the
    reference's FlipDriver calls them (flip_app_sim.cc:204, 259-274): frame 0
    is the initial state, frame 1 the state after the first frame, equal to
    the output= frame of the same run; debug=1 adds the debug dumps."""
    n = 32
    out_dir = str(tmp_path / "out")
    prefix = str(tmp_path / "frames_")
    r = subprocess.run([APP, f"n={n}", "scene=sphere_drop", "frames=1", "seed=1",
                        f"output_dir={out_dir}", f"output={prefix}", "debug=1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, _, f0 = read_frame(os.path.join(out_dir, "particles_00000000"))
    _, _, f1 = read_frame(os.path.join(out_dir, "particles_00000001"))
    dims = (n, n, n)
    mask = scenes.sphere_drop_mask(dims, 8, (16, 23, 16), 5)
    assert f0.shape[0] == int(mask.sum()) * 8
    with open(os.path.join(out_dir, "particles_00000001"), "rb") as a, \
            open(prefix + "00000001", "rb") as b:
        assert a.read() == b.read()
    # One step: the move uses the previous G2P's grid velocity (0 at first),
    # so positions are unchanged and velocities are not.
    assert not np.array_equal(f0["vel"], f1["vel"])
    dbg = os.path.join(out_dir, "debug")
    for step in ("00000000", "00000001"):
        for name in (f"particles_{step}", f"grid_velocity_{step}",
                     f"grid_velocity_{step}.txt", f"grid_phi_{step}.txt"):
            assert os.path.getsize(os.path.join(dbg, name)) > 0, name
    with open(os.path.join(dbg, "particles_00000001")) as fh:
        head = [fh.readline() for _ in range(4)]
    assert head[0].split() == ["0", "0", "0"] and int(head[2]) == f1.shape[0]
    assert head[3].startswith("0 : ")
