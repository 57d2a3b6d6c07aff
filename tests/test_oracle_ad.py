"""Pins of the oracle's implicit momentum solve (oracle/or_ad.c):
set_solve_ad, calc_residual_ad, the line sweeps and the GCM loop of
solve_ad.  CPU only: closed forms on linear and quadratic fields, and a
dense solve of the system assembled from the residual and the ghost rows."""
import numpy as np
import pytest

import cases
from flow_structure_interaction_fem_ibm_amd import pp
from oracle_lib import Oracle

incI = [0, 0, 1, 1, 0, 0, 1, 1]
incJ = [0, 1, 1, 0, 0, 1, 1, 0]
incK = [0, 0, 0, 0, 1, 1, 1, 1]


def grid(nx, ny, nz, uniform=False):
    if uniform:
        return np.linspace(0, 1, nx + 1), np.linspace(0, 0.8, ny + 1), np.linspace(0, 1.2, nz + 1)
    return pp.tanh_nodes(nx + 1, 1.0, 1.3), pp.tanh_nodes(ny + 1, 0.8, 0.8), pp.tanh_nodes(nz + 1, 1.2, 1.6)


def no_body(nx, ny, nz, uniform=False):
    xn, yn, zn = grid(nx, ny, nz, uniform)
    o = Oracle(xn, yn, zn)
    o.set_iblank(np.zeros(o.shape, np.int8))
    return o


def centres(o):
    xc, yc, zc = (o.metrics(d)[1] for d in range(3))
    return xc[None, None, :], yc[None, :, None], zc[:, None, None]


def full(o, a):
    return np.ascontiguousarray(np.broadcast_to(a, o.shape), dtype=np.float64)


def test_set_solve_ad_uniform_closed_form():
    nx, ny, nz = 9, 8, 7
    o = no_body(nx, ny, nz, uniform=True)
    h = (1.0 / nx, 0.8 / ny, 1.2 / nz)
    nu, nub, dt = 0.013, 0.021, 0.4
    visc = full(o, nu)
    bv = full(o, nub)
    coef = o.set_solve_ad(dt, visc, bv, bv, bv)
    I = (slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1))
    for d in range(3):
        am, ap = coef[2 * d][I], coef[2 * d + 1][I]
        inner = -0.5 * dt * nu / h[d] ** 2
        outer = -0.5 * dt * nub * 2.0 / h[d] ** 2
        ax = 2 - d  # numpy axis of direction d
        sl_lo = [slice(None)] * 3
        sl_lo[ax] = 0
        sl_hi = [slice(None)] * 3
        sl_hi[ax] = -1
        sl_in = [slice(None)] * 3
        sl_in[ax] = slice(1, None)
        # west/south/back coefficient: the outer face on the first cell
        assert np.allclose(am[tuple(sl_lo)], outer, rtol=1e-13)
        assert np.allclose(am[tuple(sl_in)], inner, rtol=1e-13)
        sl_in[ax] = slice(0, -1)
        assert np.allclose(ap[tuple(sl_hi)], outer, rtol=1e-13)
        assert np.allclose(ap[tuple(sl_in)], inner, rtol=1e-13)
    # nothing written outside the cells
    mask = np.ones(o.shape, bool)
    mask[I] = False
    assert not coef[:, mask].any()


@pytest.mark.parametrize("half", [True, False])
def test_calc_residual_ad_linear_fields_stretched(half):
    """var and nu linear in x, y, z, constant face velocities, stretched
    grid: the residual of r = var + half_dt (U bx + V by + W bz)
    - dt/2 (bx nux + by nuy + bz nuz) vanishes off the outer faces."""
    nx, ny, nz = 12, 10, 9
    o = no_body(nx, ny, nz)
    X, Y, Z = centres(o)
    b = (0.7, -1.3, 0.45)
    g = (0.02, 0.011, -0.015)
    var = full(o, 0.3 + b[0] * X + b[1] * Y + b[2] * Z)
    visc = full(o, 0.05 + g[0] * X + g[1] * Y + g[2] * Z)
    dt = 0.37
    coef = o.set_solve_ad(dt, visc, visc, visc, visc)
    U, V, W = 0.8, -0.55, 1.1
    faces = [full(o, U), full(o, V), full(o, W)]
    hdt = 0.5 * dt if half else 0.0
    r = var + hdt * (U * b[0] + V * b[1] + W * b[2]) - 0.5 * dt * (b[0] * g[0] + b[1] * g[1] + b[2] * g[2])
    m, loc, res = o.calc_residual_ad(hdt, coef, faces, var, r)
    J = (slice(2, nz), slice(2, ny), slice(2, nx))
    assert np.abs(res[J]).max() < 1e-13, np.abs(res[J]).max()
    # a dropped advection term would leave hdt U bx on every cell
    if half:
        bad = r - hdt * U * b[0]
        m2, _, res2 = o.calc_residual_ad(hdt, coef, faces, var, bad)
        assert np.allclose(res2[J], -hdt * U * b[0], rtol=1e-9)
    # the outer faces carry the boundary terms that belong to the right-hand
    # side (rhs_advec_diff), so the residual there is not zero
    assert m >= np.abs(res[J]).max()
    assert np.abs(res).max() == m
    i, j, k = loc
    assert abs(res[k, j, i]) == m


def test_calc_residual_ad_quadratic_uniform_diffusion_scale():
    nx, ny, nz = 10, 9, 8
    o = no_body(nx, ny, nz, uniform=True)
    X, Y, Z = centres(o)
    var = full(o, X ** 2 + 2.0 * Y ** 2 - Z ** 2)
    nu, dt = 0.09, 0.6
    coef = o.set_solve_ad(dt, full(o, nu), full(o, nu), full(o, nu), full(o, nu))
    zero = [full(o, 0.0)] * 3
    # A var = var - dt/2 nu lap(var); lap = 2 + 4 - 2 exactly on a uniform grid
    r = var - 0.5 * dt * nu * 4.0
    _, _, res = o.calc_residual_ad(0.0, coef, zero, var, r)
    J = (slice(2, nz), slice(2, ny), slice(2, nx))
    assert np.abs(res[J]).max() < 1e-13


def body_case(n=(12, 10, 9), seed=5):
    nx, ny, nz = n
    xn, yn, zn = grid(nx, ny, nz)
    ib, fr, gs, ex = pp.build_geometry(xn, yn, zn, [pp.Ellipsoid((0.48, 0.41, 0.62), (0.2, 0.17, 0.23), 20.0, 24)])
    assert gs.n > 0
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    o.set_vel_ghosts(ex["coeffD"], 2.0, 1.0)
    rng = np.random.default_rng(seed)
    return o, ib, gs, ex, rng


def dense_system(o, ib, gs, coeffD, hdt, coef, faces, vbi_c, bw=2.0, iw=1.0):
    """The converged state of solve_ad for one velocity component as a
    dense system over the cells: fluid rows from calc_residual_ad applied to
    unit fields, ghost rows u_g + iw sum_{r != self} cD_r u_r = (bw - iw
    cD_self) u_BI, solid rows u = 0.  Returns (A, rhs_ghost, cells)."""
    nz3, ny3, nx3 = o.shape
    cells = [(i, j, k) for k in range(1, nz3 - 2) for j in range(1, ny3 - 2) for i in range(1, nx3 - 2)]
    pos = {c: q for q, c in enumerate(cells)}
    N = len(cells)
    A = np.zeros((N, N))
    zero_r = np.zeros(o.shape)
    for q, (i, j, k) in enumerate(cells):
        e = np.zeros(o.shape)
        e[k, j, i] = 1.0
        _, _, res = o.calc_residual_ad(hdt, coef, faces, e, zero_r)
        col = -np.array([res[kk, jj, ii] for (ii, jj, kk) in cells])
        A[:, q] = col
    bg = np.zeros(N)
    gset = {(int(gs.ig[n]), int(gs.jg[n]), int(gs.kg[n])): n for n in range(gs.n)}
    for (i, j, k), q in pos.items():
        if ib[k, j, i] == 0:
            continue
        A[q, :] = 0.0
        A[q, q] = 1.0
        n = gset.get((i, j, k))
        if n is None:
            continue
        rhs = bw * vbi_c[n]
        for r in range(8):
            c = (int(gs.ic[n]) + incI[r], int(gs.jc[n]) + incJ[r], int(gs.kc[n]) + incK[r])
            if c == (i, j, k):
                rhs -= iw * coeffD[n, r] * vbi_c[n]
            else:
                A[q, pos[c]] += iw * coeffD[n, r]
        bg[q] = rhs
    return A, bg, cells


@pytest.mark.parametrize("half", [True, False])
def test_solve_ad_matches_dense_solve(half):
    o, ib, gs, ex, rng = body_case()
    X, Y, Z = centres(o)
    visc = full(o, 0.02 + 0.01 * X)
    dt = 0.05
    coef = o.set_solve_ad(dt, visc, visc, visc, visc)
    faces = [rng.standard_normal(o.shape) * 0.5 for _ in range(3)]
    hdt = 0.5 * dt if half else 0.0
    nl = [cases.random_field(o.shape, 10 + q) for q in range(3)]
    vbi = rng.standard_normal((gs.n, 3))
    u0 = [np.zeros(o.shape) for _ in range(3)]
    u, v, w, it, mr = o.solve_ad(hdt, coef, faces, *u0, *nl, vbi=vbi, itermax=500, restol=1e-13)
    assert mr <= 1e-13 and 1 < it < 500, (it, mr)
    A, bg, cells = dense_system(o, ib, gs, ex["coeffD"], hdt, coef, faces, vbi[:, 0])
    for comp, (got, r) in enumerate(zip((u, v, w), nl)):
        A, bg, cells = dense_system(o, ib, gs, ex["coeffD"], hdt, coef, faces, vbi[:, comp])
        b = np.array([r[k, j, i] if ib[k, j, i] == 0 else 0.0 for (i, j, k) in cells]) + bg
        x = np.linalg.solve(A, b)
        g = np.array([got[k, j, i] for (i, j, k) in cells])
        assert np.allclose(g, x, rtol=0, atol=1e-11 * np.abs(x).max()), (comp, np.abs(g - x).max())


def test_itsolv_ad_single_sweep_reduces_residual_and_keeps_solid_cells():
    o, ib, gs, ex, rng = body_case(seed=9)
    dt = 0.05
    visc = full(o, 0.03)
    coef = o.set_solve_ad(dt, visc, visc, visc, visc)
    faces = [full(o, 0.0)] * 3
    r = cases.random_field(o.shape, 21)
    var = rng.standard_normal(o.shape)
    m0, _, _ = o.calc_residual_ad(0.0, coef, faces, var, r)
    v1 = o.itsolv_ad(0.0, coef, faces, var, r)
    m1, _, _ = o.calc_residual_ad(0.0, coef, faces, v1, r)
    assert m1 < 0.1 * m0
    g = np.zeros(o.shape, bool)
    g[gs.kg, gs.jg, gs.ig] = True
    solid = (ib == 1) & ~g
    solid[0, :, :] = solid[-2:, :, :] = False
    solid[:, 0, :] = solid[:, -2:, :] = False
    solid[:, :, 0] = solid[:, :, -2:] = False
    assert np.all(v1[solid] == 0.0)
    # ghost cells keep their value through the line sweeps
    assert np.array_equal(v1[g], var[g])
