"""Water sources (AddContributionFromSources / DeleteFluidInSources,
flip_simulation.cc:1206-1277) on the GPU against oracle/oracle.py's numpy
restatement: emitted particles bit for bit (same Philox stream), region
deletion as the same multiset, the capacity error, and fsl_water_app running
an init-file scene with a source."""
import os
import subprocess

import numpy as np
import pytest

import fsl
from fsl import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

APP = os.path.join(os.path.dirname(fsl.__file__), os.pardir, "fsl_water_app")
DIMS = (32, 32, 32)


def make_ctx(capacity=200000):
    ctx = fsl.Context(DIMS, capacity=capacity)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    mask = scenes.dam_break_mask(DIMS, 8, 12)
    ctx.seed_particles(mask, 8, 5)
    return ctx


def rows(a):
    """Rows of a (n, k) array in lexicographic order."""
    a = np.asarray(a)
    return a[np.lexsort(a.T[::-1])]


@pytest.mark.parametrize("lo,hi,ppc,seed,stream", [
    ((10.0, 20.0, 10.0), (14.0, 24.0, 14.0), 8, 3, 0),       # whole cells + end layer
    ((3.4, 27.6, 0.0), (9.9, 31.5, 31.0), 6, 17, 5),         # partial cells, clipped
    ((-2.0, -1.5, 30.2), (1.2, 0.7, 40.0), 8, 1, 2),         # clipped at both ends
])
def test_source_particles_match_oracle(lo, hi, ppc, seed, stream):
    ctx = make_ctx()
    n0 = ctx.num_particles()
    pos0, vel0, gv0 = ctx.get_particles()
    vel = (0.25, -1.5, 0.0)
    ctx.add_source(lo, hi, vel, ppc, seed, stream)
    pos, v, gv = ctx.get_particles()
    ref, refv = orc.source_particles(DIMS, lo, hi, vel, ppc, seed, stream)
    print(lo, hi, "emitted", pos.shape[0] - n0, "expected", ref.shape[0])
    assert ref.shape[0] > 0
    assert pos.shape[0] == n0 + ref.shape[0]
    assert np.array_equal(pos[:n0], pos0) and np.array_equal(v[:n0], vel0)
    assert np.array_equal(pos[n0:], ref)
    assert np.array_equal(v[n0:], refv) and np.array_equal(gv[n0:], refv)
    ctx.close()


def test_delete_in_boxes_matches_oracle():
    ctx = make_ctx()
    ctx.add_source((4.0, 20.0, 4.0), (12.0, 28.0, 12.0), (0.0, -1.0, 0.0), 8, 9, 0)
    ctx.move_particles(0.0)  # sort the appended particles into their cells
    pos0, vel0, _ = ctx.get_particles()
    boxes = np.float32([[4.5, 20.0, 4.0, 9.25, 27.0, 12.0],
                        [0.0, 0.0, 0.0, 3.0, 6.5, 31.0]])
    ctx.delete_particles_in_boxes(boxes)
    pos, vel, _ = ctx.get_particles()
    gone = orc.inside_boxes(pos0, boxes)
    print("particles", pos0.shape[0], "deleted", int(gone.sum()))
    assert 0 < gone.sum() < pos0.shape[0]
    assert pos.shape[0] == pos0.shape[0] - gone.sum()
    a = np.concatenate([pos, vel], 1)
    b = np.concatenate([pos0[~gone], vel0[~gone]], 1)
    assert np.array_equal(rows(a), rows(b))
    ctx.close()


def test_source_capacity_error():
    ctx = make_ctx(capacity=40000)
    n0 = ctx.num_particles()
    with pytest.raises(fsl.FslError) as e:
        ctx.add_source((0.0, 0.0, 0.0), (31.0, 31.0, 31.0), (0, 0, 0), 8, 1, 0)
    assert e.value.code == -4  # FSL_ERR_CAPACITY
    assert ctx.num_particles() == n0
    ctx.close()


def test_water_app_init_file_with_source(tmp_path):
    init = tmp_path / "scene.txt"
    init.write_text("fluid 2\nboundary 0\ngravity 9.8\n"
                    "source-static 0.4, 0.7, 0.4, 0.6, 0.8, 0.6, 0, -5, 0\n"
                    "solid 0.8 0.0 0.8 0.9 0.2 0.9\n")
    out = str(tmp_path / "particles_")
    r = subprocess.run([APP, "n=32", f"init={init}", "frames=2", "seed=1", f"output={out}"],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    print("\n".join(lines))
    import json
    fr = [json.loads(l) for l in lines]
    assert len(fr) == 2 and all(f["steps"] >= 1 for f in fr)
    # fluid 2 at n = 32 is the cube [0, 7] x [0, 23] x [0, 31] (cell centres
    # tested): 7 x 23 x 31 cells of 8 particles. Each step the source's
    # particles that left its box (moving down) stay, so the count grows.
    assert fr[0]["particles"] > 7 * 23 * 31 * 8
    assert fr[1]["particles"] > fr[0]["particles"]
    raw = open(out + "00000002", "rb").read()
    rec = np.frombuffer(raw[28:], dtype=np.dtype([("id", "<i4"), ("pos", "<f4", 3),
                                                  ("vel", "<f4", 3)]))
    assert (rec["pos"] >= 0).all() and (rec["pos"] <= 32).all()
