"""GPU adjoint: fsl_project_backward and the checkpointed rollout gradient.

Checks
  * projection: the dot-product identity <g, P u> = <P^T g, u> with the GPU's
    own forward, and P^T g against the FP64 oracle's explicit projection
    (columns P e_f computed by the oracle built with double arithmetic);
  * rollout: replay of the recorded substeps reproduces the forward bit for
    bit; the gradients w.r.t. the initial grid velocity and the body force
    match central finite differences (eps 1e-3) of the FP64 oracle run on the
    same recorded particle positions, within 1e-4 relative;
  * the 128^3 sphere drop (8 substeps): gradients against central finite
    differences of the GPU's own replay. Under fixed positions the loss is an
    exact quadratic in (u0, force), so a central difference has no truncation
    error at any step size; steps of half the inputs' size (4.9 on the body
    force, 0.25 on a u0 of scale 0.5) keep the FP32 rounding of the loss far
    below the 1e-4 bar.
"""
import numpy as np
import pytest

import fsl
from fsl import scenes
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TIGHT = dict(precision=fsl.PCG_FP64, stop_mode=fsl.STOP_REL, threshold=1e-12,
             max_iterations=20000)


def tight():
    return fsl.solve_params(**TIGHT)


def forward_labels(n, mask, seed, ppc=8):
    """Run one step's stages up to the divergence; return the marker grid."""
    pos, vel = scenes.seed_particles(mask, ppc, seed)
    ctx = fsl.Context(n, capacity=pos.shape[0] + 64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_particles(pos, vel)
    ctx.move_particles(1.0 / 30)
    ctx.clear(); ctx.p2g(); ctx.apply_bc()
    m = ctx.get_int(fsl.GRID_MARKER)
    ctx.close()
    return m


def projection_ctx(n, marker):
    ctx = fsl.Context(n, capacity=64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_marker(marker)
    return ctx


def project(ctx, u):
    ctx.set_vec(fsl.GRID_VELOCITY, u, np.ones(ctx.box, np.uint8))
    ctx.project_forward(tight())
    return ctx.get_vec(fsl.GRID_VELOCITY)[0]


def test_projection_adjoint_dot_product():
    n = (32, 32, 32)
    marker = forward_labels(n, scenes.sphere_drop_mask(n, 8, (16, 23, 16), 5), 5)
    ctx = projection_ctx(n, marker)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(ctx.box + (3,)).astype(np.float32)
    g = rng.standard_normal(ctx.box + (3,)).astype(np.float32)
    pu = project(ctx, u)
    ptg, st = ctx.project_backward(tight(), g)
    assert st["true_rel_residual"] <= 1e-9
    lhs = float(np.dot(g.ravel().astype(np.float64), pu.ravel().astype(np.float64)))
    rhs = float(np.dot(ptg.ravel().astype(np.float64), u.ravel().astype(np.float64)))
    scale = np.linalg.norm(g) * np.linalg.norm(u)
    print(f"<g,Pu> {lhs:.8e} <P^T g,u> {rhs:.8e}")
    assert abs(lhs - rhs) <= 1e-6 * scale
    # The projection is idempotent, so P^T is too.
    ptg2, _ = ctx.project_backward(tight(), ptg)
    assert np.linalg.norm(ptg2 - ptg) <= 1e-5 * np.linalg.norm(ptg)
    ctx.close()


def test_projection_adjoint_matches_fp64_oracle():
    n = (16, 16, 8)
    # Pool plus a block: every FLUID cell has an AIR path (no sealed pocket).
    mask = np.zeros(n, np.uint8)
    mask[:, :3, :] = 1
    mask[3:9, 3:9, 2:6] = 1
    marker = forward_labels(n, mask, 7)
    ctx = projection_ctx(n, marker)
    rng = np.random.default_rng(1)
    g = rng.standard_normal(ctx.box + (3,)).astype(np.float32)
    ctx.divergence()  # labels of the system (the forward pass computes them)
    ptg, _ = ctx.project_backward(tight(), g)
    ctx.close()
    # Explicit columns of P from the FP64 oracle: only faces of FLUID cells
    # can change; (P^T g)_f = <g, P e_f>.
    o = Oracle(n, fp64=True)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    o.set_marker(marker)
    act = np.ones(o.box, np.uint8)
    zero = np.zeros(o.box)
    fl = marker == 1
    touch = np.zeros(o.box + (3,), bool)
    for d in range(3):
        touch[..., d] |= fl
        sh = np.zeros_like(fl)
        sl = [slice(None)] * 3
        sl2 = [slice(None)] * 3
        sl[d] = slice(1, None)
        sl2[d] = slice(None, -1)
        sh[tuple(sl)] = fl[tuple(sl2)]
        touch[..., d] |= sh
    ref = g.astype(np.float64).copy()
    g64 = g.astype(np.float64)
    for idx in np.argwhere(touch):
        e = np.zeros(o.box + (3,))
        e[tuple(idx)] = 1.0
        o.set_vec(0, e, act)
        o.divergence()
        o.set_scalar(1, zero)  # x0 = 0: no warm start from the previous column
        rc, _, _ = o.solve(precision=1, stop_mode=1, threshold=1e-14, max_iter=10000)
        assert rc == 0
        o.pressure_update()
        pe = o.get_vec(0)[0]
        ref[tuple(idx)] = float(np.sum(g64 * pe))
    err = np.linalg.norm(ptg - ref) / np.linalg.norm(ref)
    print(f"P^T g vs FP64 oracle columns: rel L2 {err:.2e} over {touch.sum()} columns")
    assert err <= 1e-5


# ---------------------------------------------------------------- rollouts
def target_box(box, lo, hi):
    m = np.zeros(box + (3,), np.float32)
    # box index = cell index + 1
    m[lo[0] + 1:hi[0] + 1, lo[1] + 1:hi[1] + 1, lo[2] + 1:hi[2] + 1, :] = 1.0
    return m


def make_rollout(n, mask, seed, ppc=8):
    pos, vel = scenes.seed_particles(mask, ppc, seed)
    ctx = fsl.Context(n, capacity=pos.shape[0] + 64)
    ctx.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    ctx.set_box_boundary()
    ctx.set_particles(pos, vel)
    return ctx, pos.shape[0]


def run_forward(ctx, steps, u0, prm, frame=1.0 / 30):
    ctx.rollout_begin(steps)
    step = 0.0
    up = None
    for k in range(steps):
        dt, step = ctx.compute_dt(1, frame, step)
        if step + 1e-5 >= frame:
            step = 0.0
        _, up = ctx.rollout_step(dt, prm, u0=u0 if k == 0 else None, want_u_post=True)
    return up


def test_rollout_replay_reproduces_forward():
    n = (32, 32, 32)
    ctx, _ = make_rollout(n, scenes.sphere_drop_mask(n, 8, (16, 23, 16), 5), 2)
    rng = np.random.default_rng(3)
    u0 = (0.3 * rng.standard_normal(ctx.box + (3,))).astype(np.float32)
    up = run_forward(ctx, 3, u0, tight())
    rep, _ = ctx.rollout_replay(tight(), u0=u0)
    assert np.array_equal(rep.view(np.uint32), up.view(np.uint32))
    ctx.close()


def oracle_rollout(n, ck, u0, force, flip=0.95, g=9.8):
    """FP64 oracle run of the recorded substeps with the recorded positions."""
    o = Oracle(n, fp64=True)
    o.set_params(g, tuple(float(x) for x in force), flip)
    o.set_box_boundary()
    vel_out = None
    K = len(ck)
    for k, (pos, vel, perm, _, dt) in enumerate(ck):
        pos64 = pos.astype(np.float64)
        v = vel.astype(np.float64) if k == 0 else vel_out[perm]
        o.set_particles(pos64, v, np.zeros_like(pos64))
        o.clear()
        o.p2g()
        if k == 0 and u0 is not None:
            vv, act = o.get_vec(0)
            on = act.astype(bool)
            vv[on] = u0[on]
            o.set_vec(0, vv, act)
        o.apply_bc(); o.save_velocities(); o.delete_outside()
        o.add_gravity(dt); o.apply_bc(); o.divergence()
        rc, _, _ = o.solve(precision=1, stop_mode=1, threshold=1e-14, max_iter=20000)
        assert rc == 0
        o.pressure_update()
        if k == K - 1:
            return o.get_vec(0)[0]
        assert o.velocity_update() == 0
        o.g2p()
        # Oracle particles are in its own order; map back to the recorded one.
        op, ov, _, _ = o.get_particles()
        key = {tuple(p): q for q, p in enumerate(op.astype(np.float32).view(np.uint32).tolist())}
        vel_out = np.zeros((pos.shape[0], 3))
        for q, p in enumerate(pos.view(np.uint32).tolist()):
            j = key.get(tuple(p))
            if j is not None:
                vel_out[q] = ov[j]


def loss_of(up, tmask):
    return float(np.sum((up.astype(np.float64) * tmask) ** 2))


def test_rollout_gradient_matches_fp64_finite_differences():
    n = (32, 32, 32)
    steps = 3
    ctx, cap = make_rollout(n, scenes.sphere_drop_mask(n, 8, (16, 23, 16), 5), 4)
    rng = np.random.default_rng(5)
    u0 = (0.3 * rng.standard_normal(ctx.box + (3,))).astype(np.float32)
    up = run_forward(ctx, steps, u0, tight())
    tmask = target_box(ctx.box, (10, 4, 10), (22, 16, 22))
    gu, gf, _ = ctx.rollout_backward(tight(), 2.0 * up * tmask)
    ck = [ctx.rollout_checkpoint(k, cap) for k in range(steps)]
    ctx.close()
    force = np.array([0.0, 9.8, 0.0])
    u64 = u0.astype(np.float64)
    L0 = loss_of(oracle_rollout(n, ck, u64, force), tmask)
    assert abs(L0 - loss_of(up, tmask)) <= 1e-5 * L0
    eps = 1e-3
    for d in range(3):
        e = np.zeros(3); e[d] = eps
        fd = (loss_of(oracle_rollout(n, ck, u64, force + e), tmask)
              - loss_of(oracle_rollout(n, ck, u64, force - e), tmask)) / (2 * eps)
        print(f"dL/df{d}: adjoint {gf[d]:.8e} FD {fd:.8e}")
        assert abs(gf[d] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(gf).max())
    for trial in range(2):
        e = rng.standard_normal(u0.shape)
        fd = (loss_of(oracle_rollout(n, ck, u64 + eps * e, force), tmask)
              - loss_of(oracle_rollout(n, ck, u64 - eps * e, force), tmask)) / (2 * eps)
        ad = float(np.sum(gu.astype(np.float64) * e))
        print(f"dL/du0 . e: adjoint {ad:.8e} FD {fd:.8e}")
        assert abs(ad - fd) <= 1e-4 * abs(fd)


def test_rollout_gradient_sphere_drop_128():
    dims, mk, seed = scenes.WORKLOADS["sphere_drop_128"]
    steps = 8
    ctx, _ = make_rollout(dims, mk(dims), seed)
    prm = fsl.solve_params(precision=fsl.PCG_FP64, stop_mode=fsl.STOP_REL, threshold=1e-11,
                           max_iterations=20000)
    rng = np.random.default_rng(6)
    u0 = (0.5 * rng.standard_normal(ctx.box + (3,))).astype(np.float32)
    up = run_forward(ctx, steps, u0, prm)
    tmask = target_box(ctx.box, (56, 56, 56), (72, 72, 72))
    gu, gf, _ = ctx.rollout_backward(prm, 2.0 * up * tmask)
    force = np.array(ctx.force)

    def L(u, f):
        return loss_of(ctx.rollout_replay(prm, u0=u, force=f)[0], tmask)

    assert abs(L(u0, force) - loss_of(up, tmask)) <= 1e-6 * loss_of(up, tmask)
    for d in range(3):
        h = 4.9
        e = np.zeros(3); e[d] = h
        fd = (L(u0, force + e) - L(u0, force - e)) / (2 * h)
        print(f"128^3 dL/df{d}: adjoint {gf[d]:.8e} FD {fd:.8e}")
        assert abs(gf[d] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(gf).max())
    # Along the gradient itself the derivative is |grad|^2 (no cancellation);
    # along random N(0, I) directions its typical size is |grad|, the scale
    # the error is measured against.
    gnorm = float(np.linalg.norm(gu.astype(np.float64)))
    dirs = [gu / np.float32(gnorm)] + [rng.standard_normal(u0.shape).astype(np.float32)
                                      for _ in range(2)]
    for t, e in enumerate(dirs):
        h = np.float32(0.25 if t else 0.25 * np.sqrt(e.size))
        fd = (L(u0 + h * e, force) - L(u0 - h * e, force)) / (2 * float(h))
        ad = float(np.sum(gu.astype(np.float64) * e))
        scale = abs(fd) if t == 0 else gnorm
        print(f"128^3 dL/du0 . e{t}: adjoint {ad:.8e} FD {fd:.8e} (scale {scale:.3e})")
        assert abs(ad - fd) <= 1e-4 * scale
    ctx.close()


def test_projection_gradient_sphere_drop_128_fp64_fd():
    """The 128^3 sphere drop's labels (one step's stages): the gradient of
    L(u) = sum over the 16^3 target box of |P u|^2 through the GPU adjoint
    (2 P^T (mask P u), FP64 solves), against central finite differences with
    eps = 1e-3 of the FP64 oracle's own projection of u on the same labels.
    L is quadratic in u, so the central difference has no truncation error:
    the check measures the GPU gradient against the FP64 reference."""
    dims, mk, seed = scenes.WORKLOADS["sphere_drop_128"]
    marker = forward_labels(dims, mk(dims), seed)
    ctx = projection_ctx(dims, marker)
    rng = np.random.default_rng(12)
    u = (0.5 * rng.standard_normal(ctx.box + (3,))).astype(np.float32)
    tmask = target_box(ctx.box, (56, 56, 56), (72, 72, 72))
    pu = project(ctx, u)
    g, st = ctx.project_backward(tight(), 2.0 * pu * tmask)
    ctx.close()
    assert st["true_rel_residual"] <= 1e-9
    g = g.astype(np.float64)
    o = Oracle(dims, fp64=True)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    o.set_marker(marker)
    act = np.ones(o.box, np.uint8)
    zero = np.zeros(o.box)

    def L(uu):
        o.set_vec(0, uu, act)
        o.divergence()
        o.set_scalar(1, zero)
        rc, _, _ = o.solve(precision=1, stop_mode=1, threshold=1e-14, max_iter=20000)
        assert rc == 0
        o.pressure_update()
        return loss_of(o.get_vec(0)[0], tmask)

    u64 = u.astype(np.float64)
    assert abs(L(u64) - loss_of(pu, tmask)) <= 1e-5 * loss_of(pu, tmask)
    gnorm = float(np.linalg.norm(g))
    eps = 1e-3
    dirs = [g / gnorm] + [rng.standard_normal(u.shape) for _ in range(2)]
    for t, e in enumerate(dirs):
        fd = (L(u64 + eps * e) - L(u64 - eps * e)) / (2 * eps)
        ad = float(np.sum(g * e))
        scale = abs(fd) if t == 0 else gnorm
        print(f"128^3 projection dL/du . e{t}: adjoint {ad:.8e} FP64 FD {fd:.8e} "
              f"(scale {scale:.3e})")
        assert abs(ad - fd) <= 1e-4 * scale


def test_autograd_projection_matches_c_abi():
    import torch

    from fsl.autograd import project

    n = (32, 32, 32)
    marker = forward_labels(n, scenes.sphere_drop_mask(n, 8, (16, 23, 16), 5), 8)
    ctx = projection_ctx(n, marker)
    rng = np.random.default_rng(9)
    # Host tensors: the C ABI takes host arrays, and the library runs on its
    # own HIP stream (torch's device runtime is not involved).
    u = torch.tensor(rng.standard_normal(ctx.box + (3,)), dtype=torch.float32,
                     requires_grad=True)
    w = torch.tensor(rng.standard_normal(ctx.box + (3,)), dtype=torch.float32)
    loss = (project(ctx, u, tight()) * w).sum()
    loss.backward()
    ref, _ = ctx.project_backward(tight(), w.cpu().numpy())
    assert np.allclose(u.grad.cpu().numpy(), ref, rtol=0, atol=1e-6 * np.abs(ref).max())
    ctx.close()


# Run in a fresh process that initializes torch's GPU runtime before the
# library's: torch's lazy CUDA initialization has failed ("No HIP GPUs are
# available") in a process where the library had already run rollouts.
_DEVICE_AUTOGRAD = r"""
import sys
import numpy as np
import torch
torch.cuda.init()
sys.path[:0] = [{root!r}, {pkg!r}, {tests!r}]
import fsl
from fsl import scenes
from fsl.autograd import project
from test_gpu_adjoint import forward_labels, projection_ctx, tight

n = (32, 32, 32)
marker = forward_labels(n, scenes.sphere_drop_mask(n, 8, (16, 23, 16), 5), 8)
ctx = projection_ctx(n, marker)
rng = np.random.default_rng(11)
u_h = rng.standard_normal(ctx.box + (3,)).astype(np.float32)
w_h = rng.standard_normal(ctx.box + (3,)).astype(np.float32)
dev = torch.device("cuda:0")
u = torch.tensor(u_h, device=dev, requires_grad=True)
w = torch.tensor(w_h, device=dev)
up = project(ctx, u, tight())
assert up.is_cuda
up_dev = up.detach().cpu().numpy()
(up * w).sum().backward()
assert u.grad.is_cuda
g_dev = u.grad.cpu().numpy()
# The same projection through host arrays.
ctx.set_vec(fsl.GRID_VELOCITY, u_h, np.ones(ctx.box, np.uint8))
ctx.project_forward(tight())
up_host, _ = ctx.get_vec(fsl.GRID_VELOCITY)
g_host, _ = ctx.project_backward(tight(), w_h)
assert np.array_equal(up_dev, up_host), "forward values differ"
assert np.array_equal(g_dev, g_host), "gradients differ"
ctx.close()
print("device autograd ok")
"""


def test_autograd_projection_device_tensors_in_place():
    """CUDA tensors cross the C ABI in place (no host staging): the forward
    values and the gradient equal the host-array path's bit for bit, and stay
    on the device."""
    import os
    import subprocess
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    script = _DEVICE_AUTOGRAD.format(root=root, pkg=os.path.join(root, "fluid-sim-lb_amd"),
                                     tests=here)
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "device autograd ok" in r.stdout


def test_projection_with_ic_matches_jacobi():
    """The block IC(0) preconditioner changes the solver's path, not the
    system: the projection and its transpose match the Jacobi ones."""
    n = (32, 32, 32)
    marker = forward_labels(n, scenes.sphere_drop_mask(n, 8, (16, 23, 16), 5), 5)
    ctx = projection_ctx(n, marker)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(ctx.box + (3,)).astype(np.float32)
    g = rng.standard_normal(ctx.box + (3,)).astype(np.float32)
    ic = dict(TIGHT, preconditioner=fsl.PRECOND_IC)
    out = {}
    for name, prm in (("jacobi", tight()), ("ic", fsl.solve_params(**ic))):
        ctx.set_vec(fsl.GRID_VELOCITY, u, np.ones(ctx.box, np.uint8))
        st_f = ctx.project_forward(prm)
        pu = ctx.get_vec(fsl.GRID_VELOCITY)[0].astype(np.float64)
        ptg, st_b = ctx.project_backward(prm, g)
        print(f"{name}: forward {st_f['iterations']} iterations, backward "
              f"{st_b['iterations']}")
        out[name] = (pu, ptg.astype(np.float64), st_f["iterations"], st_b["iterations"])
    ctx.close()
    (pj, gj, fj, bj), (pi, gi, fi, bi) = out["jacobi"], out["ic"]
    assert fi < fj and bi < bj
    assert np.linalg.norm(pi - pj) <= 1e-6 * np.linalg.norm(pj)
    assert np.linalg.norm(gi - gj) <= 1e-6 * np.linalg.norm(gj)
