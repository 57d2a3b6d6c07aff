"""Pins of the oracle (the CPU restatement of the reference path) against
closed forms, brute force and an independently assembled sparse operator.
No GPU needed."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import cases
from oracle_lib import Oracle, tdma


def test_make_grid_uniform_closed_form():
    # CORE_GRID.F90:40-44,97-149 on a uniform grid: x(i) = (i-1)h, centres,
    # widths h, centre spacing h everywhere (dxc(1) = h via x(0) = -x(2))
    n, L = 16, 2.0
    o = Oracle(cases.uniform(n, L), cases.uniform(8), cases.uniform(8))
    x, xc, dx, dxinv, dxc, dxcinv = o.metrics(0)
    h = L / n
    assert np.allclose(x[1:n + 2], np.arange(n + 1) * h, rtol=0, atol=1e-14)
    assert x[0] == -x[2]
    assert np.allclose(dx[: n + 1], h, atol=1e-14)
    assert np.allclose(xc[: n + 1], (np.arange(n + 1) - 0.5) * h, atol=1e-14)
    assert np.allclose(dxc[1 : n + 1], h, atol=1e-14)
    assert np.allclose(dxcinv[1 : n + 1] * dxinv[1 : n + 1], 1 / h**2, rtol=1e-13)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tdma_matches_dense_solve(seed):
    rng = np.random.default_rng(seed)
    n = 37
    a, c = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 2, n)
    r = rng.normal(size=n)
    M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    u = tdma(a, b, c, r)
    assert np.allclose(u, np.linalg.solve(M, r), rtol=1e-13, atol=1e-13)


def test_tdma_constant_coefficient_closed_form():
    # -u'' = 0 with u(0)=0, u(n+1)=1 discretised: u_i = i/(n+1)
    n = 20
    a = np.ones(n); b = -2 * np.ones(n); c = np.ones(n); r = np.zeros(n)
    r[-1] = -1.0
    u = tdma(a, b, c, r)
    assert np.allclose(u, np.arange(1, n + 1) / (n + 1), atol=1e-14)


def test_residual_matches_independent_assembly():
    xn, yn, zn, ib, gs, _ = cases.sphere_case(14, r=0.22, stretch=1.5)
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    A, unknown = cases.assemble(xn, yn, zn, ib, gs)
    p = cases.random_field(ib.shape, 3)
    b = cases.random_field(ib.shape, 4, ib)
    res, resm, loc, l2 = o.calc_residual(p, b)
    ref = (b.ravel() - A @ p.ravel()) * (1 - ib.ravel())
    assert np.allclose(res.ravel(), ref, rtol=1e-12, atol=1e-10)
    assert resm == pytest.approx(np.abs(ref).max(), rel=1e-13)
    k, j, i = np.unravel_index(np.argmax(np.abs(res)), res.shape)
    assert loc == (i, j, k)
    assert l2 == pytest.approx(np.linalg.norm(ref), rel=1e-12)


def test_residual_second_order_truncation():
    # manufactured p = cos(pi x)cos(pi y)cos(pi z): dp/dn = 0 on [0,1]^3,
    # lap p = -3 pi^2 p; the discrete residual at cell centres is O(h^2)
    errs = []
    for n in (16, 32):
        xn = cases.uniform(n)
        o = Oracle(xn, xn, xn)
        o.set_iblank(np.zeros((n + 3,) * 3, np.int8))
        xc = o.metrics(0)[1]
        X, Y, Z = np.meshgrid(xc[: n + 3], xc[: n + 3], xc[: n + 3], indexing="ij")
        p = (np.cos(np.pi * X) * np.cos(np.pi * Y) * np.cos(np.pi * Z)).transpose(2, 1, 0).copy()
        b = -3 * np.pi**2 * p
        _, resm, _, _ = o.calc_residual(p, b)
        errs.append(resm)
    assert 3.3 < errs[0] / errs[1] < 4.7


def test_mean_remove_closed_form():
    xn, yn, zn, ib, gs, _ = cases.sphere_case(12, r=0.25, stretch=1.0)
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    p = cases.random_field(ib.shape, 7) + 3.0
    V = cases.volumes(xn, yn, zn)
    fl = (1 - ib) * (V > 0)
    expect = (p * V * fl).sum() / (V * fl).sum()
    mean = o.mean_remove(p)
    assert mean == pytest.approx(expect, rel=1e-13)
    assert abs((p * V * fl).sum()) < 1e-13 * (V * fl).sum()
    assert np.all(cases.interior(p)[cases.interior(ib) == 1] == 0.0)


def test_ghost_marking_bruteforce():
    rng = np.random.default_rng(5)
    n = 10
    xn = cases.uniform(n)
    ib = np.zeros((n + 3,) * 3, np.int8)
    ib[1:-2, 1:-2, 1:-2] = (rng.uniform(size=(n, n, n)) < 0.4)
    o = Oracle(xn, xn, xn)
    ng = o.set_iblank(ib)
    gm = o.flags()["ghost"]
    expect = np.zeros_like(gm)
    for k in range(1, n + 1):
        for j in range(1, n + 1):
            for i in range(1, n + 1):
                if ib[k, j, i] != 1:
                    continue
                nb = [(k, j, max(i - 1, 1)), (k, j, min(i + 1, n)), (k, max(j - 1, 1), i),
                      (k, min(j + 1, n), i), (max(k - 1, 1), j, i), (min(k + 1, n), j, i)]
                if any(ib[t] == 0 for t in nb):
                    expect[k, j, i] = 1
    assert np.array_equal(gm, expect)
    assert ng == expect.sum()
    ig, jg, kg = o.ghost_list()
    order = np.lexsort((ig, jg, kg))
    assert np.array_equal(order, np.arange(ng))   # i fastest, then j, then k


def test_ghost_rows_fixed_point_and_constants():
    xn, yn, zn, ib, gs, _ = cases.sphere_case(16, r=0.2, stretch=1.2)
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    # constant field on fluid cells: every ghost becomes the constant (sum c_r = 1)
    p = np.where(ib == 1, 0.0, 2.5)
    p[0, :, :] = p[:, 0, :] = p[:, :, 0] = 2.5
    o.outer_ghost(p)
    o.ghost_rows(p)
    gidx = (gs.kg, gs.jg, gs.ig)
    assert np.allclose(p[gidx], 2.5, rtol=0, atol=1e-13)
    # random field: converged ghost values satisfy the ghost rows
    p = cases.random_field(ib.shape, 9, ib)
    o.outer_ghost(p)
    o.ghost_rows(p)
    A, _ = cases.assemble(xn, yn, zn, ib, gs)
    r = A @ p.ravel()
    gflat = gs.ig + (xn.size + 2) * (gs.jg + (yn.size + 2) * gs.kg)
    assert np.abs(r[gflat]).max() < 1e-13


@pytest.mark.parametrize("omega", [1.0, 0.7])
def test_point_jacobi_matches_assembled_operator(omega):
    """One MG_itsolv_Point_Jacobi sweep is var + omega D^-1 (b - A var) on
    fluid rows (A assembled independently), holds ghost cells and scales the
    other blank cells by 1 - omega."""
    xn, yn, zn, ib, gs, _ = cases.sphere_case(14, r=0.22, stretch=1.5)
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    A, unknown = cases.assemble(xn, yn, zn, ib, gs)
    p = cases.random_field(ib.shape, 5)
    b = cases.random_field(ib.shape, 6, ib)
    out = o.point_jacobi(p, b, omega).ravel()
    pf, bf = p.ravel(), b.ravel()
    shp = ib.shape
    inner = np.zeros(shp, bool)
    inner[1:-2, 1:-2, 1:-2] = True
    inner = inner.ravel()
    fl = inner & (ib.ravel() == 0)
    ghost = np.zeros(pf.size, bool)
    ghost[(gs.ig + shp[2] * (gs.jg + shp[1] * gs.kg)).astype(np.int64)] = True
    solid = inner & (ib.ravel() == 1) & ~ghost
    d = A.diagonal()
    ref = pf + omega * (bf - A @ pf) / np.where(fl, d, 1.0)
    assert np.allclose(out[fl], ref[fl], rtol=1e-12, atol=1e-12)
    assert np.array_equal(out[ghost], pf[ghost])
    assert np.allclose(out[solid], (1 - omega) * pf[solid], rtol=1e-15, atol=0)
    assert np.array_equal(out[~inner], pf[~inner])


def test_apply_transpose_is_adjoint():
    xn, yn, zn, ib, gs, _ = cases.sphere_case(12, r=0.25, stretch=1.0)
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    A, unknown = cases.assemble(xn, yn, zn, ib, gs)
    u = unknown.reshape(ib.shape)
    x = cases.random_field(ib.shape, 1) * u
    y = cases.random_field(ib.shape, 2) * u
    Ax = o.apply_A(x)
    ATy = o.apply_AT(y)
    assert np.allclose((A @ x.ravel()) * unknown, Ax.ravel(), rtol=1e-12, atol=1e-9)
    assert (y * Ax).sum() == pytest.approx((x * ATy).sum(), rel=1e-11)


def left_null_vector(A, unknown, V):
    """w with A^T w = 0 on the unknowns, normalised so that the sum of w over
    the fluid cells equals the fluid volume (V: volumes, zero on blank cells)."""
    idx = np.flatnonzero(unknown)
    As = A[idx][:, idx].tocsc()
    n = len(idx)
    import scipy.sparse as sp
    # bordered system [A^T e; f^T 0] [w; s] = [0; sum V_f], f = fluid indicator
    v = V.ravel()[idx]
    f = (v > 0).astype(float)
    e = np.ones(n)
    M = sp.bmat([[As.T, sp.csc_matrix(e[:, None])], [sp.csc_matrix(f[None, :]), None]]).tocsc()
    rhs = np.zeros(n + 1)
    rhs[-1] = v.sum()
    sol = spla.spsolve(M, rhs)
    w = np.zeros(A.shape[0])
    w[idx] = sol[:n]
    return w


def direct_solution(A, unknown, b, V, ib):
    """Zero-mean solution of A p = b (b made consistent first) by sparse LU."""
    import scipy.sparse as sp
    idx = np.flatnonzero(unknown)
    As = A[idx][:, idx].tocsc()
    n = len(idx)
    v = (V.ravel() * (1 - ib.ravel()))[idx]
    M = sp.bmat([[As, sp.csc_matrix(np.ones((n, 1)))], [sp.csc_matrix(v[None, :]), None]]).tocsc()
    sol = spla.spsolve(M, np.concatenate([b.ravel()[idx], [0.0]]))
    p = np.zeros(A.shape[0])
    p[idx] = sol[:n]
    return p.reshape(b.shape), sol[n]


def test_oracle_mg_converges_to_direct_solution():
    xn, yn, zn, ib, gs, _ = cases.sphere_case(16, r=0.2, stretch=1.0)
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    A, unknown = cases.assemble(xn, yn, zn, ib, gs)
    V = cases.volumes(xn, yn, zn)
    w = left_null_vector(A, unknown, V)
    fluid = (unknown.reshape(ib.shape)) & (ib == 0)
    b = cases.random_field(ib.shape, 11, ib)
    b[fluid] -= (w.reshape(ib.shape)[fluid] * b[fluid]).sum() / w.reshape(ib.shape)[fluid].sum()
    pd, lam = direct_solution(A, unknown, b, V, ib)
    assert abs(lam) < 1e-10 * np.abs(b).max()     # consistent after projection
    p = np.zeros(ib.shape)
    rc, st = o.mg_solve(p, b, rel_tol=1e-13, iter_max=300)
    assert rc == 0, st
    o.mean_remove(p)
    err = np.linalg.norm((p - pd)[fluid]) / np.linalg.norm(pd[fluid])
    assert err < 1e-10, err
