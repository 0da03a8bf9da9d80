"""The adjoint formula of the projection, checked on the CPU.

The projection u' = P u (divergence, Jacobi-PCG solve, pressure update) has
the transpose P^T g = g + G~ A^-1 D (M g): M keeps the faces the pressure
update writes, D is the divergence on FLUID rows, A the assembled Laplacian
and G~ the unmasked pressure difference. Here that formula, written out in
numpy/scipy in FP64, is compared with the explicit columns P e_f of the FP64
oracle (the oracle's projection built with double arithmetic). The GPU
kernels implement the same formula.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.oracle import Oracle

AIR, FLUID, SOLID = 0, 1, 2


def labels(n, seed):
    """Box-format marker: solid ring, FLUID pool plus random FLUID cells."""
    rng = np.random.default_rng(seed)
    box = tuple(d + 2 for d in n)
    m = np.zeros(box, np.int32)
    m[0, :, :] = m[-1, :, :] = SOLID
    m[:, 0, :] = m[:, -1, :] = SOLID
    m[:, :, 0] = m[:, :, -1] = SOLID
    inner = m[1:-1, 1:-1, 1:-1]
    inner[:, :2, :] = FLUID
    extra = rng.random(inner.shape) < 0.25
    inner[extra] = FLUID
    inner[3, 4, 3] = SOLID  # an interior solid cell next to fluid
    return m


def adjoint_formula(marker, g):
    box = marker.shape
    row = np.full(box, -1, np.int64)
    inner = np.zeros(box, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    rows = np.argwhere((marker == FLUID) & inner)
    for r, c in enumerate(rows):
        row[tuple(c)] = r
    nr = len(rows)
    ns = marker != SOLID
    # A (solver.cc:197-233) and rhs = D (M g).
    I, J, V = [], [], []
    rhs = np.zeros(nr)
    for r, c in enumerate(rows):
        diag = 0
        for d in range(3):
            for s in (-1, 1):
                nb = c.copy(); nb[d] += s
                t = tuple(nb)
                if marker[t] in (FLUID, AIR):
                    diag += 1
                if row[t] >= 0:
                    I.append(r); J.append(row[t]); V.append(-1.0)
            up = c.copy(); up[d] += 1
            dn = c.copy(); dn[d] -= 1
            rhs[r] += (g[tuple(up) + (d,)] if ns[tuple(up)] else 0.0) \
                - (g[tuple(c) + (d,)] if ns[tuple(dn)] else 0.0)
        I.append(r); J.append(r); V.append(float(diag))
    A = sp.csr_matrix((V, (I, J)), shape=(nr, nr))
    y = spla.spsolve(A.tocsc(), rhs)
    yb = np.zeros(box)
    for r, c in enumerate(rows):
        yb[tuple(c)] = y[r]
    out = g.copy()
    for d in range(3):
        dn = np.zeros(box)
        sl = [slice(None)] * 3; sl2 = [slice(None)] * 3
        sl[d] = slice(1, None); sl2[d] = slice(None, -1)
        dn[tuple(sl)] = yb[tuple(sl2)]
        out[..., d] += yb - dn
    return out


def test_adjoint_formula_matches_fp64_oracle_columns():
    n = (8, 8, 8)
    marker = labels(n, 0)
    rng = np.random.default_rng(1)
    box = marker.shape
    g = rng.standard_normal(box + (3,))
    ptg = adjoint_formula(marker, g)
    o = Oracle(n, fp64=True)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    o.set_marker(marker)
    act = np.ones(box, np.uint8)
    zero = np.zeros(box)
    ref = np.empty_like(g)
    for idx in np.ndindex(*box, 3):
        e = np.zeros(box + (3,))
        e[idx] = 1.0
        o.set_vec(0, e, act)
        o.divergence()
        o.set_scalar(1, zero)  # x0 = 0: no warm start from the previous column
        rc, _, _ = o.solve(precision=1, stop_mode=1, threshold=1e-14, max_iter=10000)
        assert rc in (0,)
        o.pressure_update()
        ref[idx] = float(np.sum(g * o.get_vec(0)[0]))
    err = np.linalg.norm(ptg - ref) / np.linalg.norm(ref)
    assert err < 1e-10, err


def test_projection_of_oracle_is_idempotent_and_divergence_free():
    n = (8, 8, 8)
    marker = labels(n, 2)
    box = marker.shape
    o = Oracle(n, fp64=True)
    o.set_params(9.8, (0.0, 9.8, 0.0), 0.95)
    o.set_box_boundary()
    o.set_marker(marker)
    act = np.ones(box, np.uint8)
    u = np.random.default_rng(3).standard_normal(box + (3,))

    def P(x):
        o.set_vec(0, x, act)
        o.divergence()
        o.set_scalar(1, np.zeros(box))
        rc, _, _ = o.solve(precision=1, stop_mode=1, threshold=1e-14, max_iter=10000)
        assert rc == 0
        o.pressure_update()
        return o.get_vec(0)[0]

    pu = P(u)
    assert np.linalg.norm(P(pu) - pu) <= 1e-10 * np.linalg.norm(pu)
    o.set_vec(0, pu, act)
    o.divergence()
    div = o.get_scalar(0)
    assert np.abs(div[marker == FLUID]).max() < 1e-10
