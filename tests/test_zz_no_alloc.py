"""fsl_get_scalar_noalloc (the allocation-free variant of fsl_get_scalar,
staged through the library's fixed device array) against fsl_get_scalar:
bit-identical pressure and divergence on two calls in a row, the second on a
smaller domain, and a clear error for a box beyond the array's size."""
import numpy as np
import pytest

import fsl
from fsl import scenes

pytestmark = pytest.mark.gpu


def stepped(n, seed):
    mask = scenes.sphere_drop_mask(n, max(2, n[1] // 4), (n[0] // 2, (2 * n[1]) // 3, n[2] // 2),
                                   min(n) // 6)
    pos, vel = scenes.seed_particles(mask, 8, seed)
    ctx = fsl.Context(n, capacity=pos.shape[0] + 64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_particles(pos, vel)
    prm = fsl.solve_params(precision=fsl.PCG_MIXED, stop_mode=fsl.STOP_REL, threshold=1e-6)
    ctx.step(1.0 / 30, prm)
    return ctx


def test_noalloc_readback_matches_allocating_one():
    for n, seed in (((48, 40, 32), 3), ((16, 16, 24), 4)):
        ctx = stepped(n, seed)
        for which in (fsl.GRID_PRESSURE, fsl.GRID_DIVERGENCE):
            a = ctx.get_scalar(which)
            b = ctx.get_scalar_noalloc(which)
            assert b.shape == a.shape
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (n, which)
        assert np.abs(ctx.get_scalar(fsl.GRID_PRESSURE)).max() > 0
        ctx.close()


def test_noalloc_rejects_a_box_beyond_its_array():
    ctx = fsl.Context((208, 208, 208), capacity=64)  # box 210^3 > 8,388,608 voxels
    with pytest.raises(fsl.FslError) as e:
        ctx.get_scalar_noalloc(fsl.GRID_PRESSURE)
    assert e.value.code == -1 and "staging array" in str(e.value)
    ctx.close()
