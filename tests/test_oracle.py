"""CPU checks that pin the oracle (no GPU).

The reference ships no tests, golden vectors or fixtures for this path, so
each oracle stage is pinned by a closed form, an invariant, a brute-force
computation or the published known answers of an algorithm it uses:

  scene seeding     Random123 Philox4x32-10 known-answer vectors
  P2G (a7)          hand-computed single-particle weights; partition of unity;
                    reproduction of a uniform field
  G2P (a24)         exact interpolation of per-component linear fields;
                    flip = 1 with dv = 0 keeps particle velocities
  Laplacian (a15)   A p == -div(grad p) (summation by parts) against a
                    brute-force dense Laplacian built from the labels
  projection        div(u_new) on FLUID cells == b - A p (the final residual),
                    i.e. ~0 after an FP64 solve to 1e-12
  PCG (a17-a21)     FP64 PCG solution == scipy dense solve of the same system
  CFL, dt (a6)      closed forms and the clamp branches
  advection (a4/5)  clamp rules and deletion in solids / outside the box
  extrapolation     two-hop corner value (order dependence) (a23)
  hydrostatics      a pool at rest stays at rest
  free fall         a blob at rest ends its first step at (0, -g dt, 0)
                    exactly (gravity, save / update, blend); with an initial
                    velocity, at (u0, v0 - g dt, w0) to rounding
  G2P blend         v + flip (dv + vel - v) on linear fields, two flips
  boundary cond.    the zeroed faces of a hand-built solid configuration
  velocity update   dv = v_new - v_saved bit for bit; the unsaved-face check
  block IC(0)       L L^T == A on A's block pattern; exact Cholesky on a chain;
                    scipy triangular solves; independent tiles
  fixtures          tests/golden/projection_systems.npz (make_golden.py): FP32
                    divergence bit-exact, dense FP64 solutions, IC(0) diagonals
"""
import os

import numpy as np
import pytest
import scipy.linalg

from fsl import scenes
from oracle.oracle import Oracle

B = 1  # box offset: box index = cell index + 1


def test_philox_known_answers():
    f = scenes.philox4x32_10
    assert [int(x[0]) for x in f([0], [0], [0], [0], 0, 0)] == [
        0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    m = 0xFFFFFFFF
    assert [int(x[0]) for x in f([m], [m], [m], [m], m, m)] == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert [int(x[0]) for x in f([0x243F6A88], [0x85A308D3], [0x13198A2E], [0x03707344],
                                 0xA4093822, 0x299F31D0)] == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_seeding_jitter_stays_in_cell():
    n = (16, 16, 16)
    mask = np.ones(n, np.uint8)
    pos, vel = scenes.seed_particles(mask, 8, 123)
    assert pos.shape == (16 ** 3 * 8, 3) and not vel.any()
    cells = np.floor(pos).astype(int)
    expect = np.repeat(np.argwhere(mask), 8, axis=0)
    assert np.array_equal(cells, expect)


def test_mask_counts_match_the_workloads():
    n = (64, 64, 64)
    assert int(scenes.WORKLOADS["sphere_drop_64"][1](n).sum()) == 69760
    n = (128, 128, 128)
    assert int(scenes.WORKLOADS["dam_break_128"][1](n).sum()) == 491520
    assert int(scenes.WORKLOADS["sphere_drop_128"][1](n).sum()) == 557840


def one_particle(pos, vel, n=(8, 8, 8)):
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), 0.95)
    o.set_particles(np.array([pos], np.float32), np.array([vel], np.float32))
    o.clear()
    o.p2g()
    return o


def test_p2g_single_particle_hand_computed():
    pos = (2.3, 3.7, 4.1)
    o = one_particle(pos, (1.0, 2.0, 3.0))
    w, act = o.get_vec(2)
    v, _ = o.get_vec(0)
    p = np.array(pos, np.float64)
    # Component d sits at the lower d-face: offset 0.5 on the other two axes.
    for d in range(3):
        lower = np.floor(p - 0.5)
        lower[d] = np.floor(p[d])
        frac = p - (lower + 0.5)
        frac[d] = p[d] - lower[d]
        for corner in np.ndindex(2, 2, 2):
            c = np.array(corner)
            wt = np.prod(np.where(c == 1, frac, 1.0 - frac))
            idx = tuple((lower.astype(int) + c) + B)
            assert abs(w[idx + (d,)] - wt) < 2e-6, (d, corner)
            assert act[idx] == 1
    # 8 corners per component; a face touched by any component is active.
    assert act.sum() <= 24 and act.sum() >= 8
    # Normalised velocity is the particle's where the weight is >= 1e-5.
    for d, val in enumerate((1.0, 2.0, 3.0)):
        m = w[..., d] >= 1e-5
        assert np.allclose(v[..., d][m], val, rtol=2e-6)


def test_p2g_partition_of_unity_and_uniform_field():
    n = (16, 16, 16)
    mask = scenes.sphere_drop_mask(n, 4, (8, 10, 8), 3)
    pos, vel = scenes.seed_particles(mask, 8, 5)
    vel[:] = (0.25, -1.5, 3.0)
    o = Oracle(n)
    o.set_particles(pos, vel)
    o.clear()
    o.p2g()
    w, _ = o.get_vec(2)
    v, act = o.get_vec(0)
    assert np.allclose(w.reshape(-1, 3).sum(axis=0), pos.shape[0], rtol=1e-5)
    for d, val in enumerate((0.25, -1.5, 3.0)):
        m = w[..., d] >= 1e-5
        assert np.allclose(v[..., d][m], val, rtol=1e-5, atol=1e-6)
        assert np.all(v[..., d][(act == 1) & ~m] == 0)


def test_g2p_exact_on_linear_fields():
    n = (16, 16, 16)
    rng = np.random.default_rng(0)
    pos = (rng.random((400, 3)) * 13 + 1.5).astype(np.float32)
    vel = rng.standard_normal((400, 3)).astype(np.float32)
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), 0.0)  # flip = 0: PIC, vel = v(x)
    o.set_particles(pos, vel)
    ii = np.indices(tuple(x + 2 for x in n)).astype(np.float64) - B
    coef = np.array([[0.5, 0.1, -0.2, 0.3], [-1.0, 0.05, 0.2, -0.1], [2.0, -0.3, 0.1, 0.15]])
    field = np.zeros(ii.shape[1:] + (3,))
    for d in range(3):
        loc = [ii[a] + (0.5 if a != d else 0.0) for a in range(3)]
        field[..., d] = coef[d, 0] + sum(coef[d, 1 + a] * loc[a] for a in range(3))
    act = np.ones(ii.shape[1:], np.uint8)
    o.set_vec(0, field.astype(np.float32), act)
    o.set_vec(1, np.zeros_like(field, np.float32), act)
    o.g2p()
    p, v, gv, _ = o.get_particles()
    # Extrapolation only touches the boundary layers; particles are interior.
    expect = coef[:, 0][None, :] + p.astype(np.float64) @ coef[:, 1:].T
    assert np.allclose(gv, expect, atol=2e-5)
    assert np.allclose(v, expect, atol=2e-5)


def test_g2p_flip_one_with_zero_update_keeps_velocity():
    n = (8, 8, 8)
    rng = np.random.default_rng(1)
    pos = (rng.random((50, 3)) * 5 + 1.5).astype(np.float32)
    vel = rng.standard_normal((50, 3)).astype(np.float32)
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), 1.0)
    o.set_particles(pos, vel)
    act = np.ones((10, 10, 10), np.uint8)
    o.set_vec(0, rng.standard_normal((10, 10, 10, 3)).astype(np.float32), act)
    o.set_vec(1, np.zeros((10, 10, 10, 3), np.float32), act)
    o.g2p()
    p, v, _, _ = o.get_particles()
    order_in = np.lexsort(pos.T[::-1])
    order_out = np.lexsort(p.T[::-1])
    assert np.allclose(v[order_out], vel[order_in], atol=1e-6)


def random_labels(n, seed, p_fluid=0.6, p_solid=0.1):
    rng = np.random.default_rng(seed)
    box = tuple(x + 2 for x in n)
    marker = np.zeros(box, np.int32)
    inner = rng.random(n)
    m = np.where(inner < p_fluid, 1, np.where(inner < p_fluid + p_solid, 2, 0))
    marker[1:-1, 1:-1, 1:-1] = m
    solid = np.zeros(box, bool)
    solid[0, :, :] = solid[-1, :, :] = solid[:, 0, :] = solid[:, -1, :] = True
    solid[:, :, 0] = solid[:, :, -1] = True
    marker[solid] = 2
    return marker


def dense_laplacian(marker):
    """Brute force: -Laplacian with Neumann SOLID and Dirichlet AIR."""
    fl = np.argwhere(marker == 1)
    index = {tuple(c): r for r, c in enumerate(fl)}
    A = np.zeros((len(fl), len(fl)))
    for r, c in enumerate(fl):
        for d in range(3):
            for s in (-1, 1):
                nb = c.copy()
                nb[d] += s
                m = marker[tuple(nb)]
                if m != 2:
                    A[r, r] += 1
                if m == 1:
                    A[r, index[tuple(nb)]] -= 1
    return A, fl


def drop_pocketless(marker):
    """Make every FLUID connected component touch AIR (else A is singular)."""
    from scipy import ndimage
    lab, k = ndimage.label(marker == 1)
    air = ndimage.binary_dilation(marker == 0)
    for comp in range(1, k + 1):
        if not np.any(air[lab == comp]):
            marker[lab == comp] = 0
    return marker


def test_laplacian_is_minus_div_grad():
    n = (8, 8, 8)
    marker = drop_pocketless(random_labels(n, 3))
    A, fl = dense_laplacian(marker)
    rng = np.random.default_rng(4)
    p = np.zeros(marker.shape, np.float32)
    pv = rng.standard_normal(len(fl)).astype(np.float32)
    p[tuple(fl.T)] = pv
    o = Oracle(n)
    o.set_marker(marker)
    o.set_scalar(1, p)
    box = marker.shape
    o.set_vec(0, np.zeros(box + (3,), np.float32), np.ones(box, np.uint8))
    o.pressure_update()
    o.divergence()
    div = o.get_scalar(0)
    assert np.allclose(div[tuple(fl.T)], -(A @ pv.astype(np.float64)), atol=1e-5)


def test_projection_is_divergence_free_and_matches_dense_solve():
    n = (8, 8, 8)
    marker = drop_pocketless(random_labels(n, 7, p_fluid=0.7))
    A, fl = dense_laplacian(marker)
    rng = np.random.default_rng(8)
    box = marker.shape
    v = rng.standard_normal(box + (3,)).astype(np.float32)
    o = Oracle(n)
    o.set_marker(marker)
    o.set_vec(0, v, np.ones(box, np.uint8))
    o.divergence()
    b = o.get_scalar(0)[tuple(fl.T)].astype(np.float64)
    rc, it, rr = o.solve(precision=1, stop_mode=1, threshold=1e-12, max_iter=5000)
    assert rc == 0 and it > 0
    x = o.solution()
    xd = scipy.linalg.solve(A, b)
    assert np.allclose(x, xd, rtol=1e-9, atol=1e-9)
    o.pressure_update()
    o.divergence()
    assert np.abs(o.get_scalar(0)[tuple(fl.T)]).max() < 1e-5


def test_fp32_pcg_stops_on_reference_rule():
    n = (16, 16, 16)
    marker = drop_pocketless(random_labels(n, 11, p_fluid=0.8, p_solid=0.05))
    rng = np.random.default_rng(12)
    box = marker.shape
    o = Oracle(n)
    o.set_marker(marker)
    o.set_vec(0, rng.standard_normal(box + (3,)).astype(np.float32), np.ones(box, np.uint8))
    o.divergence()
    thr = float(np.float32(np.float32(16 ** 3) * np.float32(0.5e-8)))
    rc, it, rr = o.solve(precision=0, stop_mode=0, threshold=thr, max_iter=1500)
    hist = o.residual_history()
    assert rc == 0 and rr < thr
    assert len(hist) == it + 1 and np.all(hist[:-1] >= thr)


def test_cfl_closed_form_and_exclusions():
    n = (8, 8, 8)
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), 0.95)
    box = (10, 10, 10)
    v = np.zeros(box + (3,), np.float32)
    act = np.ones(box, np.uint8)
    o.set_vec(0, v, act)
    assert o.cfl() == np.float32(1.0 / np.float32(np.sqrt(0.5 * np.float64(np.float32(9.8)))))
    v[3 + B, 7 + B, 2 + B, 0] = 50.0   # u at j == end: excluded
    v[8 + B, 2 + B, 2 + B, 1] = 60.0   # outside the partition box: excluded
    v[2 + B, 2 + B, 2 + B, 2] = -7.0   # counted
    o.set_vec(0, v, act)
    assert o.cfl() == np.float32(1.0 / 7.0)


def test_dt_clamps():
    n = (8, 8, 8)
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), 0.95)
    fs = np.float32(1.0 / 30)
    dt, step = o.compute_dt(1, fs, 0.0)       # CFL dt 0.45 > frame: one step
    assert dt == fs and step == fs
    box = (10, 10, 10)
    v = np.zeros(box + (3,), np.float32)
    v[2, 2, 2, 0] = 1000.0                     # dt below frame/10: floored
    o.set_vec(0, v, np.ones(box, np.uint8))
    dt, step = o.compute_dt(1, fs, 0.0)
    assert dt == np.float32(fs / np.float32(10))
    dt, step = o.compute_dt(1, fs, np.float32(0.029))  # snap to the frame end
    assert dt == np.float32(fs - np.float32(0.029))


def test_advection_clamps_and_deletes():
    n = (8, 8, 8)
    pos = np.array([[0.5, 0.5, 0.5], [7.5, 3.5, 3.5], [3.5, 7.5, 3.5]], np.float32)
    gv = np.array([[-30, 0, 0], [30, 0, 0], [0, 30, 0]], np.float32)
    for tank, survivors in ((False, 1), (True, 2)):
        o = Oracle(n)
        o.set_box_boundary(tank)
        o.set_particles(pos, np.zeros_like(pos), gv)
        o.move_particles(1.0)
        p, _, _, cell = o.get_particles()
        assert p.shape[0] == survivors
        assert p[0, 0] == 0.0 and cell[0, 0] == 0        # clamped to 0
        if tank:                                          # top cell kept, at y = n
            assert p[1, 1] == 8.0 and cell[1, 1] == 8
            o.clear(); o.p2g(); o.delete_outside()
            assert o.num_particles() == 1


def test_extrapolation_two_hop_corner():
    n = (8, 8, 8)
    o = Oracle(n)
    box = (10, 10, 10)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(box + (3,)).astype(np.float32)
    o.set_vec(0, v, np.ones(box, np.uint8))
    o.set_vec(1, v, np.ones(box, np.uint8))
    o.extrapolate()
    e, _ = o.get_vec(0)
    # x-upper copies (n-1,0,k) -> (n,0,k); y-lower then copies (n,0,k) -> (n,-1,k).
    k = 3
    assert e[8 + B, -1 + B, k + B, 2] == v[7 + B, 0 + B, k + B, 2]
    # Corner (-1,-1,k) is outside every copy range: untouched.
    assert e[-1 + B, -1 + B, k + B, 2] == v[-1 + B, -1 + B, k + B, 2]


def test_row_index_is_lexicographic():
    n = (8, 8, 8)
    marker = drop_pocketless(random_labels(n, 21))
    o = Oracle(n)
    o.set_marker(marker)
    o.set_vec(0, np.zeros(marker.shape + (3,), np.float32), np.ones(marker.shape, np.uint8))
    o.divergence()
    o.solve(precision=1, stop_mode=1, threshold=1e-8)
    idx = o.get_int(2)
    fl = marker == 1
    assert np.array_equal(idx[fl], np.arange(fl.sum()))
    assert np.all(idx[~fl] == -1)


def test_hydrostatic_pool_stays_at_rest():
    n = (16, 16, 16)
    mask = np.zeros(n, np.uint8)
    mask[:, :6, :] = 1
    pos, vel = scenes.seed_particles(mask, 8, 2)
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), 0.95)
    o.set_box_boundary()
    o.set_particles(pos, vel)
    step = 0.0
    for _ in range(3):
        dt, step = o.compute_dt(1, 1.0 / 30, step)
        rc, _ = o.run_one_step(dt, precision=1, stop_mode=1, threshold=1e-10, max_iter=3000)
        assert rc == 0
    _, v, _, _ = o.get_particles()
    # Without pressure the pool would fall at 3 g dt = 0.98 after 3 steps;
    # gravity is balanced by pressure, and the residual motion is the free
    # surface's sub-cell roughness.
    assert np.abs(v).max() < 0.2 * 9.8 / 30
    assert np.median(np.abs(v)) < 1e-3 * 9.8 / 30


def free_fall_blob(vel0, flip=0.95):
    """A cube of fluid (cells 12..19 on every axis, 8 particles per cell) in
    a 32^3 box, far from every wall, after one step of dt = 1/30."""
    n = (32, 32, 32)
    mask = np.zeros(n, np.uint8)
    mask[12:20, 12:20, 12:20] = 1
    pos, vel = scenes.seed_particles(mask, 8, 5)
    vel[:] = vel0
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), flip)
    o.set_box_boundary()
    o.set_particles(pos, vel)
    dt = np.float32(1.0 / 30)
    rc, _ = o.run_one_step(dt, precision=1, stop_mode=1, threshold=1e-12, max_iter=2000)
    assert rc == 0
    return o, pos, dt


def test_free_fall_at_rest_closed_form():
    """AddGravity (flip_simulation.cc:505-515), SaveGridVelocities /
    GetVelocityUpdate (:490-496, :780-801) and the FLIP blend (:1187): a blob
    at rest, far from the walls, falls freely during its first step. P2G
    gives 0 on every active face, gravity -g dt, the divergence of that
    uniform field is 0 on every FLUID cell, so the pressure is 0, dv = -g dt,
    and v + flip (dv + vel - v) = -g dt. Away from the blob's surface (where
    the G2P stencil also reads inactive faces) every particle ends with
    velocity (0, fl(0 - fl(dt g)), 0) exactly, at its old position. A flipped
    gravity sign, a dropped dt, dv = v_new (no subtraction) with a
    nonzero saved velocity (next test) or a misplaced flip term fail."""
    o, pos0, dt = free_fall_blob((0.0, 0.0, 0.0))
    p, v, gv, _ = o.get_particles()
    assert p.shape == pos0.shape
    g_dt = np.float32(0.0) - np.float32(dt * np.float32(9.8))
    assert g_dt < 0
    cell = np.floor(p).astype(int)
    inner = np.all((cell >= 13) & (cell <= 18), axis=1)
    assert inner.sum() == 6 ** 3 * 8
    assert np.array_equal(v[:, 0], np.zeros(len(v), np.float32))
    assert np.array_equal(v[:, 2], np.zeros(len(v), np.float32))
    assert np.array_equal(v[inner, 1], np.full(inner.sum(), g_dt, np.float32))
    assert np.array_equal(gv[inner, 1], np.full(inner.sum(), g_dt, np.float32))
    # Particles sat still during the step's advection (grid velocity 0).
    order_in = np.lexsort(pos0.T[::-1])
    order_out = np.lexsort(p.T[::-1])
    assert np.array_equal(p[order_out], pos0[order_in])
    # Nothing anywhere moved upwards or sideways.
    assert (v[:, 1] <= 0).all() and (v[:, 1] >= g_dt).all()


def test_free_fall_with_initial_velocity_closed_form():
    """The same blob moving at (u0, v0, w0): the saved velocity is
    (u0, v0, w0), the step's update dv = v_new - v_saved = (0, -g dt, 0), and
    the particles end at (u0, v0 - g dt, w0) to rounding (P2G's weighted
    average of equal values rounds). With dv = v_new instead, the blend would
    add 0.95 (u0, v0, w0) again; with PIC only (flip ignored) the result would
    be the same here, so the next test pins the blend itself."""
    u = np.array([0.7, -1.3, 0.4], np.float32)
    o, pos0, dt = free_fall_blob(tuple(u))
    p, v, _, _ = o.get_particles()
    cell = np.floor(p).astype(int)
    inner = np.all((cell >= 14) & (cell <= 17), axis=1)
    assert inner.sum() > 0
    expect = np.array([u[0], u[1] - dt * np.float32(9.8), u[2]], np.float64)
    err = np.abs(v[inner].astype(np.float64) - expect).max()
    assert err <= 1e-5, err
    # The particles moved by u dt during the step's advection (grid velocity
    # = initial velocity).
    assert np.allclose(np.sort(p[:, 0]), np.sort(pos0[:, 0] + u[0] * dt), atol=1e-5)


def test_g2p_flip_blend_closed_form():
    """UpdateSingleParticleFromGrid's blend (flip_simulation.cc:1187):
    vel' = v(x) + flip (dv(x) + vel - v(x)), with v and dv linear fields (the
    trilinear interpolation reproduces them to rounding) and arbitrary
    particle velocities; flip = 0.95 and flip = 0.3. A swapped sign on dv or
    on v inside the bracket, or the blend weights exchanged, fail."""
    n = (16, 16, 16)
    rng = np.random.default_rng(7)
    pos = (rng.random((300, 3)) * 12 + 2.0).astype(np.float32)
    vel = rng.standard_normal((300, 3)).astype(np.float32)
    ii = np.indices(tuple(x + 2 for x in n)).astype(np.float64) - B
    cv = np.array([[0.5, 0.1, -0.2, 0.3], [-1.0, 0.05, 0.2, -0.1], [2.0, -0.3, 0.1, 0.15]])
    cu = np.array([[-0.2, 0.02, 0.01, -0.03], [0.3, -0.01, 0.04, 0.0], [0.1, 0.0, -0.02, 0.05]])
    fv = np.zeros(ii.shape[1:] + (3,))
    fu = np.zeros(ii.shape[1:] + (3,))
    for d in range(3):
        loc = [ii[a] + (0.5 if a != d else 0.0) for a in range(3)]
        fv[..., d] = cv[d, 0] + sum(cv[d, 1 + a] * loc[a] for a in range(3))
        fu[..., d] = cu[d, 0] + sum(cu[d, 1 + a] * loc[a] for a in range(3))
    act = np.ones(ii.shape[1:], np.uint8)
    for flip in (0.95, 0.3):
        o = Oracle(n)
        o.set_params(9.8, (0, 9.8, 0), flip)
        o.set_particles(pos, vel)
        o.set_vec(0, fv.astype(np.float32), act)
        o.set_vec(1, fu.astype(np.float32), act)
        o.g2p()
        p, v, _, _ = o.get_particles()
        order_in = np.lexsort(pos.T[::-1])
        order_out = np.lexsort(p.T[::-1])
        x = p[order_out].astype(np.float64)
        vx = cv[:, 0][None, :] + x @ cv[:, 1:].T
        ux = cu[:, 0][None, :] + x @ cu[:, 1:].T
        expect = vx + flip * (ux + vel[order_in].astype(np.float64) - vx)
        assert np.abs(v[order_out] - expect).max() <= 3e-5, flip


def test_boundary_conditions_hand_listed_faces():
    """ApplyBoundaryConditions (flip_simulation.cc:656-713) on a hand-built
    case: solid cells (3,3,3) and (4,3,3) inside an 8^3 box, faces active
    (value 1 in every component) on the voxels of [2,4]^3 only. A solid cell
    zeroes the normal component on each face it shares with a non-solid
    neighbour: its lower faces are stored at the cell, its upper faces at
    the neighbour (c + e_d); the face between the two solid cells is kept;
    an inactive face (the upper x face of (4,3,3), stored at (5,3,3)) stays
    inactive. The expected zeros are listed by hand."""
    n = (8, 8, 8)
    o = Oracle(n)
    o.set_params(9.8, (0, 9.8, 0), 0.95)
    o.set_box_boundary()
    o.add_solid_box((3, 3, 3), (4, 3, 3))
    box = tuple(x + 2 for x in n)
    act = np.zeros(box, np.uint8)
    act[2 + B:5 + B, 2 + B:5 + B, 2 + B:5 + B] = 1
    vals = np.where(act[..., None] == 1, 1.0, 0.0).astype(np.float32) * np.ones(3, np.float32)
    o.set_vec(0, vals, act)
    o.apply_bc()
    v, a = o.get_vec(0)
    assert np.array_equal(a, act)
    zeros = {((3, 3, 3), 0), ((3, 3, 3), 1), ((3, 3, 3), 2),  # lower faces of (3,3,3)
             ((3, 4, 3), 1), ((3, 3, 4), 2),                  # its upper y and z faces
             ((4, 3, 3), 1), ((4, 3, 3), 2),                  # lower y, z faces of (4,3,3)
             ((4, 4, 3), 1), ((4, 3, 4), 2)}                  # its upper y and z faces
    for i in range(2, 5):
        for j in range(2, 5):
            for k in range(2, 5):
                for d in range(3):
                    want = 0.0 if ((i, j, k), d) in zeros else 1.0
                    assert v[i + B, j + B, k + B, d] == want, ((i, j, k), d)
    m = o.get_int(0)
    assert m[3 + B, 3 + B, 3 + B] == 2 and m[4 + B, 3 + B, 3 + B] == 2
    assert m[2 + B, 3 + B, 3 + B] != 2


def test_velocity_update_is_new_minus_saved():
    """SaveGridVelocities then GetVelocityUpdate (flip_simulation.cc:490-496,
    780-801): dv = v_new - v_saved on every face active in both, one FP32
    subtraction each (so bit-equal to numpy's); a face active in v but never
    saved is the reference's CHECK failure."""
    n = (8, 8, 8)
    rng = np.random.default_rng(3)
    box = tuple(x + 2 for x in n)
    act = (rng.random(box) < 0.7).astype(np.uint8)
    act[0] = act[:, 0] = act[:, :, 0] = 0  # the update loops over [0, n]
    v1 = rng.standard_normal(box + (3,)).astype(np.float32) * act[..., None]
    v2 = rng.standard_normal(box + (3,)).astype(np.float32) * act[..., None]
    o = Oracle(n)
    o.set_vec(0, v1, act)
    o.save_velocities()
    o.set_vec(0, v2, act)
    assert o.velocity_update() == 0
    du, da = o.get_vec(1)
    on = act == 1
    assert np.array_equal(da, act)
    assert np.array_equal(du[on].view(np.uint32), (v2[on] - v1[on]).view(np.uint32))
    # A face active in v but not saved.
    extra = act.copy()
    idx = tuple(np.argwhere(act[1:, 1:, 1:] == 0)[0] + 1)
    extra[idx] = 1
    o.set_vec(0, v2, extra)
    assert o.velocity_update() != 0


def test_sealed_pocket_and_enclosed_cell_errors():
    n = (8, 8, 8)
    box = (10, 10, 10)
    marker = np.full(box, 2, np.int32)
    marker[3:6, 3:6, 3:6] = 1       # sealed pocket with no AIR: singular
    o = Oracle(n)
    o.set_marker(marker)
    o.set_vec(0, np.random.default_rng(0).standard_normal(box + (3,)).astype(np.float32),
              np.ones(box, np.uint8))
    o.divergence()
    rc, _, _ = o.solve(precision=0, stop_mode=0, threshold=1e-30, max_iter=200)
    # Singular (pure Neumann) system: the solve returns a status, never aborts.
    assert rc in (0, -3, -4, -5)
    marker = np.full(box, 2, np.int32)
    marker[4, 4, 4] = 1             # one FLUID cell enclosed by SOLID: zero diagonal
    o2 = Oracle(n)
    o2.set_marker(marker)
    o2.set_vec(0, np.zeros(box + (3,), np.float32), np.ones(box, np.uint8))
    o2.divergence()
    rc, _, _ = o2.solve()
    assert rc == -2


# ---- water sources (oracle/oracle.py's numpy restatement)
from oracle import oracle as _orc  # noqa: E402


def test_oracle_philox_known_answers():
    f = _orc.philox
    assert [int(x[0]) for x in f([0], [0], [0], [0], 0, 0)] == [
        0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    m = 0xFFFFFFFF
    assert [int(x[0]) for x in f([m], [m], [m], [m], m, m)] == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert [int(x[0]) for x in f([0x243F6A88], [0x85A308D3], [0x13198A2E], [0x03707344],
                                 0xA4093822, 0x299F31D0)] == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def _source_scalar(dims, lo, hi, ppc, seed, stream):
    """AddContributionFromSources' loops (flip_simulation.cc:1222-1262) one
    cell and one attempt at a time."""
    lo = [np.float32(v) for v in lo]
    hi = [np.float32(v) for v in hi]
    out = []
    for i in range(max(int(lo[0]), 0), min(int(hi[0]), dims[0] - 1) + 1):
        for j in range(max(int(lo[1]), 0), min(int(hi[1]), dims[1] - 1) + 1):
            for k in range(max(int(lo[2]), 0), min(int(hi[2]), dims[2] - 1) + 1):
                lin = (i * dims[1] + j) * dims[2] + k
                kept = 0
                for a in range(32):
                    if kept >= ppc:
                        break
                    x = [int(v[0]) for v in _orc.philox([lin], [a], [2], [0], seed, stream)]
                    p = [np.float32(c) + np.float32((x[e] >> 8) / 16777216.0)
                         for e, c in enumerate((i, j, k))]
                    if all(lo[e] <= p[e] <= hi[e] for e in range(3)):
                        out.append(p)
                        kept += 1
    return np.array(out, np.float32).reshape(-1, 3)


def test_source_particles_match_scalar_loops():
    dims = (12, 10, 9)
    lo, hi = (2.3, 0.0, 6.5), (5.7, 2.25, 9.8)  # clipped in z, partial cells
    pos, vel = _orc.source_particles(dims, lo, hi, (0.5, -1.0, 0.0), 6, 11, 3)
    ref = _source_scalar(dims, lo, hi, 6, 11, 3)
    assert pos.shape[0] > 0
    assert np.array_equal(pos, ref)
    assert np.array_equal(vel, np.tile(np.float32([0.5, -1.0, 0.0]), (pos.shape[0], 1)))


def test_source_particles_closed_form_counts():
    # The box as Source::get_box returns it for a scaled cube [0, 5]^3: end
    # reduced to 4. Cells 0..3 lie inside (every attempt kept: 8 each); a
    # point of a cell at index 4 is inside only at u = 0 on each axis.
    dims = (16, 16, 16)
    pos, _ = _orc.source_particles(dims, (0, 0, 0), (4, 4, 4), (0, 0, 0), 8, 7, 0)
    assert pos.shape[0] == 4 ** 3 * 8
    cells = np.floor(pos).astype(int)
    assert cells.min() == 0 and cells.max() == 3
    _, counts = np.unique(cells, axis=0, return_counts=True)
    assert (counts == 8).all() and counts.size == 64
    # Another stream gives other points in the same cells.
    pos2, _ = _orc.source_particles(dims, (0, 0, 0), (4, 4, 4), (0, 0, 0), 8, 7, 1)
    assert pos2.shape == pos.shape and not np.array_equal(pos, pos2)


def test_inside_boxes_bounds_inclusive():
    pos = np.float32([[1, 1, 1], [2, 2, 2], [2.0001, 1, 1], [0.5, 0.5, 0.5]])
    m = _orc.inside_boxes(pos, [[1, 1, 1, 2, 2, 2]])
    assert m.tolist() == [True, True, False, False]
    m2 = _orc.inside_boxes(pos, [[1, 1, 1, 2, 2, 2], [0, 0, 0, 0.5, 0.5, 0.5]])
    assert m2.tolist() == [True, True, False, True]


# ---- block IC(0) (the IC branch of solver.cc:103-106, openvdb.patch:209-224) ----

def _block_pattern(A, fl, n, tiles):
    """A with the couplings between cells of different tiles dropped."""
    w = np.array(n) // np.array(tiles)
    t = [tuple((c - 1) // w) for c in fl]  # box index - 1 = cell index
    Ab = A.copy()
    for i, j in zip(*np.nonzero(A)):
        if t[i] != t[j]:
            Ab[i, j] = 0.0
    return Ab


def _l_from_diag(Ab, d):
    """The IC(0) factor of a 7-point Laplacian from its diagonal: every lower
    entry of column m is A_im / L_mm = -1 / L_mm."""
    L = np.tril(Ab, -1) / d[None, :]
    L[np.diag_indices_from(L)] = d
    return L


@pytest.mark.parametrize("tiles", [(1, 1, 1), (2, 2, 1), (1, 2, 2)])
def test_ic0_factor_reproduces_a_on_its_pattern(tiles):
    n = (8, 8, 8)
    marker = drop_pocketless(random_labels(n, 21, p_fluid=0.75, p_solid=0.05))
    A, fl = dense_laplacian(marker)
    o = Oracle(n)
    o.set_marker(marker)
    d, _ = o.ic_factor(tiles)
    assert d.shape == (len(fl),) and np.all(d > 0)
    Ab = _block_pattern(A, fl, n, tiles)
    L = _l_from_diag(Ab, d.astype(np.float64))
    M = L @ L.T
    pat = Ab != 0
    # IC(0) defining property: L L^T equals A wherever A has an entry.
    assert np.abs(M[pat] - Ab[pat]).max() < 1e-5
    # and the factor differs from A only by fill outside the pattern.
    assert np.abs((M - Ab)[pat]).max() < 1e-5 and np.abs(M - Ab).max() > 0


def test_ic0_is_exact_cholesky_on_a_chain():
    n = (4, 4, 16)
    box = tuple(x + 2 for x in n)
    marker = np.zeros(box, np.int32)
    marker[0, :, :] = marker[-1, :, :] = marker[:, 0, :] = marker[:, -1, :] = 2
    marker[:, :, 0] = marker[:, :, -1] = 2
    marker[2, 2, 1:-1] = 1  # one FLUID column in AIR: a tridiagonal system
    A, fl = dense_laplacian(marker)
    o = Oracle(n)
    o.set_marker(marker)
    d, _ = o.ic_factor()
    chol = np.linalg.cholesky(A)
    assert np.allclose(d, np.diag(chol), rtol=2e-7, atol=0)


@pytest.mark.parametrize("tiles", [(1, 1, 1), (2, 1, 2)])
def test_ic0_apply_solves_the_factored_system(tiles):
    n = (8, 8, 8)
    marker = drop_pocketless(random_labels(n, 23, p_fluid=0.8, p_solid=0.05))
    A, fl = dense_laplacian(marker)
    rng = np.random.default_rng(24)
    r = rng.standard_normal(len(fl)).astype(np.float32).astype(np.float64)
    o = Oracle(n)
    o.set_marker(marker)
    d, z = o.ic_factor(tiles, r)
    L = _l_from_diag(_block_pattern(A, fl, n, tiles), d)
    y = scipy.linalg.solve_triangular(L, r, lower=True)
    zd = scipy.linalg.solve_triangular(L.T, y, lower=False)
    assert np.linalg.norm(z - zd) / np.linalg.norm(zd) < 1e-5


def test_ic0_tile_blocks_are_independent():
    n = (8, 8, 8)
    marker = drop_pocketless(random_labels(n, 25, p_fluid=0.8, p_solid=0.05))
    _, fl = dense_laplacian(marker)
    o = Oracle(n)
    o.set_marker(marker)
    d1, _ = o.ic_factor((1, 1, 1))
    d2, _ = o.ic_factor((2, 1, 1))
    x = fl[:, 0] - 1
    # Rows of the lower tile come first and see no later row: same bits.
    assert np.array_equal(d1[x < 4], d2[x < 4])
    # The first layer of the upper tile loses its x-1 coupling: its pivots
    # grow back towards A_ii.
    first = (x == 4)
    assert np.all(d2[first] >= d1[first]) and np.any(d2[first] > d1[first])


def test_ic_pcg_matches_dense_solve_in_fewer_iterations():
    n = (8, 8, 8)
    marker = drop_pocketless(random_labels(n, 27, p_fluid=0.8, p_solid=0.05))
    A, fl = dense_laplacian(marker)
    rng = np.random.default_rng(28)
    box = marker.shape
    v = rng.standard_normal(box + (3,)).astype(np.float32)
    its = {}
    for pc in (0, 1):
        o = Oracle(n)
        o.set_marker(marker)
        o.set_vec(0, v, np.ones(box, np.uint8))
        o.divergence()
        b = o.get_scalar(0)[tuple(fl.T)].astype(np.float64)
        rc, it, rr = o.solve(precision=1, stop_mode=1, threshold=1e-12, max_iter=5000, precond=pc)
        assert rc == 0
        its[pc] = it
        assert np.allclose(o.solution(), scipy.linalg.solve(A, b), rtol=1e-9, atol=1e-9)
    assert its[1] < its[0]


# ---- committed fixtures (tests/golden/make_golden.py) ----

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "projection_systems.npz")


def golden_system(s):
    g = np.load(GOLDEN)
    return {k: g[f"{k}_{s}"] for k in ("dims", "tiles", "marker", "faces", "b", "x", "ic",
                                        "ic_tiles")}


@pytest.mark.parametrize("s", range(3))
def test_oracle_matches_golden_systems(s):
    g = golden_system(s)
    n = tuple(int(d) for d in g["dims"])
    rows = g["marker"] == 1
    o = Oracle(n)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    o.set_marker(g["marker"].astype(np.int32))
    o.set_vec(0, g["faces"], np.ones(g["marker"].shape, np.uint8))
    o.apply_bc()
    o.divergence()
    assert np.array_equal(o.get_scalar(0)[rows], g["b"])
    for tiles, key in (((1, 1, 1), "ic"), (tuple(int(t) for t in g["tiles"]), "ic_tiles")):
        d, _ = o.ic_factor(tiles)
        assert np.max(np.abs(d - g[key]) / g[key]) < 1e-6
    rc, it, _ = o.solve(precision=1, stop_mode=1, threshold=1e-13, max_iter=5000)
    assert rc == 0
    x = o.solution()
    assert np.linalg.norm(x - g["x"]) / np.linalg.norm(g["x"]) < 1e-9
