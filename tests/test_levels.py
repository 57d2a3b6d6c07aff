"""CPU pins of the numpy restatement of the multigrid levels (tests/levels.py),
which the GPU tests of the single-level kernels compare with: closed forms
on uniform grids, the level-0 residual against the independently assembled
sparse operator, constant and fluid-volume-mean restriction, exact linear
reproduction of the prolongation, and the exact solution as a fixed point of
the line pass."""
import numpy as np
import pytest

import cases
from flow_structure_interaction_fem_ibm_amd import pp
from levels import FL_BLANK, Hierarchy, dir_metrics


def sphere(nx, ny, nz, r=0.22, stretch=1.3):
    xn = pp.tanh_nodes(nx, 1.0, stretch)
    yn = pp.tanh_nodes(ny, 1.0, stretch * 0.8)
    zn = pp.tanh_nodes(nz, 1.0, stretch * 1.1)
    ib, fr, gs, ex = pp.build_geometry(xn, yn, zn, [pp.Ellipsoid((0.5, 0.47, 0.53), (r, r, r), 0.0, 48)])
    ghost = np.zeros(ib.shape, bool)
    ghost[gs.kg, gs.jg, gs.ig] = True
    return xn, yn, zn, ib, gs, ghost


def test_hierarchy_sizes():
    z = np.zeros
    H = Hierarchy(np.linspace(0, 4, 257), np.linspace(0, 3, 193), np.linspace(0, 3, 193), z((195, 195, 259)))
    assert H.dims == [(256, 192, 192), (128, 96, 96), (64, 48, 48), (32, 24, 24), (16, 12, 12)]
    H = Hierarchy(*(np.linspace(0, 1, 65),) * 3, z((67, 67, 67)))
    assert H.dims == [(64, 64, 64), (32, 32, 32), (16, 16, 16)]
    # odd counts halve too: the last coarse cell takes three fine cells
    H = Hierarchy(np.linspace(0, 1, 129), np.linspace(0, 1, 49), np.linspace(0, 1, 19), z((21, 51, 131)))
    assert H.dims == [(128, 48, 18), (64, 24, 9), (32, 12, 4)]
    assert H.nodes[2][2][-1] == 1.0 and len(H.nodes[2][2]) == 5
    H = Hierarchy(np.linspace(0, 1, 131), np.linspace(0, 1, 99), np.linspace(0, 1, 83), z((85, 101, 133)))
    assert H.dims == [(130, 98, 82), (65, 49, 41), (32, 24, 20), (16, 12, 10)]
    # a direction under 4 cells stays
    H = Hierarchy(np.linspace(0, 1, 129), np.linspace(0, 1, 97), np.linspace(0, 1, 7), z((9, 99, 131)))
    assert H.dims == [(128, 96, 6), (64, 48, 3), (32, 24, 3)] and H.cf[2] == (2, 2, 1)


def test_uniform_metrics_closed_form():
    n, h = 32, 1.0 / 32
    H = Hierarchy(*(np.linspace(0, 1, n + 1),) * 3, np.zeros((n + 3,) * 3))
    for l in range(H.nlev):
        hl = h * 2 ** l
        for d in range(3):
            m = H.metrics[l][d]
            assert np.allclose(m["w"], hl, rtol=1e-14)
            assert np.allclose(m["cm"], 1.0 / hl ** 2, rtol=1e-12)
            assert np.allclose(m["cp"], 1.0 / hl ** 2, rtol=1e-12)
            assert np.allclose(m["xc"], (np.arange(n >> l) + 0.5) * hl, rtol=1e-14)


def test_stretched_metrics_follow_make_grid():
    """On the finest level the metrics are make_grid's (CORE_GRID.F90:115-149)
    as tests/cases.py restates them: cm = 1/dxc(i)/dx(i), cp = 1/dxc(i+1)/dx(i)."""
    xs = pp.tanh_nodes(40, 2.0, 1.7)
    x, xc, dx, dxc = cases.metrics(xs)
    m = dir_metrics(xs)
    n = len(xs) - 1
    idx = np.arange(1, n + 1)
    assert np.allclose(m["w"], dx[idx], rtol=1e-15)
    assert np.allclose(m["cm"], 1.0 / dxc[idx] / dx[idx], rtol=1e-13)
    assert np.allclose(m["cp"][:-1], 1.0 / dxc[idx[:-1] + 1] / dx[idx[:-1]], rtol=1e-13)


def test_coarse_blank_and_face_flags_brute_force():
    xn, yn, zn, ib, gs, ghost = sphere(32, 24, 28)
    H = Hierarchy(xn, yn, zn, ib, ghost)
    for l in range(1, H.nlev):
        fx, fy, fz = H.cf[l]
        f = H.flags[l]
        cz, cy, cx = f.shape
        for K in range(cz):
            for J in range(cy):
                for I in range(cx):
                    ch = H.blank[l - 1][K * fz:(K + 1) * fz, J * fy:(J + 1) * fy, I * fx:(I + 1) * fx]
                    assert bool(f[K, J, I] & FL_BLANK) == bool(ch.all())
        # a fluid cell next to a blank one has that face closed; blank cells
        # keep only the outer faces
        b = H.blank[l]
        for (K, J, I) in zip(*np.nonzero(~b)):
            for bit, (dk, dj, di) in ((4, (0, 0, -1)), (8, (0, 0, 1)), (16, (0, -1, 0)), (32, (0, 1, 0)),
                                      (64, (-1, 0, 0)), (128, (1, 0, 0))):
                k, j, i = K + dk, J + dj, I + di
                outer = not (0 <= k < cz and 0 <= j < cy and 0 <= i < cx)
                assert bool(f[K, J, I] & bit) == (outer or bool(b[k, j, i]))
    # level 0: outer faces only, ghost marks from the ghost list
    f0 = H.flags[0]
    assert ((f0 & 2) != 0).sum() == gs.n
    inner = f0[1:-1, 1:-1, 1:-1]
    assert not (inner & 0xFC).any()


def test_level0_residual_matches_the_assembled_operator():
    xn, yn, zn, ib, gs, ghost = sphere(20, 18, 16)
    H = Hierarchy(xn, yn, zn, ib, ghost)
    A, unknown = cases.assemble(xn, yn, zn, ib, gs)
    rng = np.random.default_rng(5)
    e = np.zeros(ib.shape)
    e[1:-2, 1:-2, 1:-2] = rng.standard_normal((16, 18, 20))
    e *= (1 - ib) + ghost          # values on fluid and ghost cells
    rhs = cases.random_field(ib.shape, 6, ib)
    ref = rhs - (A @ e.reshape(-1)).reshape(ib.shape)
    fl = ib[1:-2, 1:-2, 1:-2] == 0
    got = H.residual(0, e[1:-2, 1:-2, 1:-2], rhs[1:-2, 1:-2, 1:-2])
    assert np.allclose(got[fl], ref[1:-2, 1:-2, 1:-2][fl], rtol=1e-12, atol=1e-9)
    assert not got[~fl].any()


@pytest.mark.parametrize("dims", [(24, 20, 16), (25, 21, 17)])
def test_restriction_constants_and_volume_mean(dims):
    xn, yn, zn, ib, gs, ghost = sphere(*dims)
    H = Hierarchy(xn, yn, zn, ib, ghost)
    one = np.ones(H.dims[0][::-1])
    c = H.restrict(0, one)
    assert np.allclose(c[~H.blank[1]], 1.0, rtol=1e-15) and not c[H.blank[1]].any()
    rng = np.random.default_rng(2)
    r = rng.standard_normal(one.shape)
    c = H.restrict(0, r)
    V = H.volumes(0)
    nz, ny, nx = one.shape
    cz, cy, cx = c.shape
    last = lambda C, n, N: slice(2 * C, n if C == N - 1 else 2 * C + 2)
    for (K, J, I) in [(0, 0, 0), (3, 4, 5), (7, 9, 11), (4, 5, 6), (cz - 1, cy - 1, cx - 1), (cz - 1, 3, cx - 1)]:
        sl = (last(K, nz, cz), last(J, ny, cy), last(I, nx, cx))
        f = ~H.blank[0][sl]
        expect = (V[sl] * r[sl])[f].sum() / V[sl][f].sum() if f.any() else 0.0
        assert np.isclose(c[K, J, I], expect, rtol=1e-14, atol=0)


@pytest.mark.parametrize("dims", [(32, 24, 28), (33, 25, 27)])
def test_prolongation_reproduces_linear_fields(dims):
    xn, yn, zn, ib, gs, ghost = sphere(*dims, r=0.12)
    H = Hierarchy(xn, yn, zn, ib, ghost)
    for l in range(H.nlev - 1):
        f_lin = lambda M, sh: (0.3 + 1.7 * M[0]["xc"][None, None, :] - 0.9 * M[1]["xc"][None, :, None]
                               + 2.3 * M[2]["xc"][:, None, None]) * np.ones(sh)
        ec = f_lin(H.metrics[l + 1], H.dims[l + 1][::-1])
        ef0 = np.zeros(H.dims[l][::-1])
        out = H.prolong(l, ec, ef0)
        exact = f_lin(H.metrics[l], ef0.shape)
        # cells whose eight partners are fluid and inside the level
        P = [H.prolong_weights(l, d) for d in range(3)]
        cfl = ~H.blank[l + 1]
        K, J, I = np.meshgrid(*(np.arange(s) for s in ef0.shape), indexing="ij")
        ok = ~H.blank[l]
        for d, G in ((0, I), (1, J), (2, K)):
            ok &= (P[d][0][G] != P[d][1][G]) | (H.cf[l + 1][d] == 1)
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    ok &= cfl[P[2][c][K], P[1][b][J], P[0][a][I]]
        assert ok.sum() > 100
        assert np.allclose(out[ok], exact[ok], rtol=1e-13, atol=1e-13)
        # constants: every fluid fine cell with a fluid partner gets the constant
        out1 = H.prolong(l, np.full(H.dims[l + 1][::-1], 2.5), ef0)
        assert np.allclose(out1[ok], 2.5, rtol=1e-15)
        assert not out1[H.blank[l]].any()


@pytest.mark.parametrize("l", [0, 1])
def test_line_pass_keeps_the_exact_solution(l):
    """The line pass is a fixed-point iteration of A_l e = rhs: an e that
    solves the level equations (blank rows hold their value) is unchanged,
    and a random e gets a smaller residual."""
    xn, yn, zn, ib, gs, ghost = sphere(24, 20, 22)
    H = Hierarchy(xn, yn, zn, ib, ghost)
    rng = np.random.default_rng(9)
    shp = H.dims[l][::-1]
    gh = (H.flags[l] & 2) != 0
    e = rng.standard_normal(shp) * (~H.blank[l] | gh)
    amx, apx, amy, apy, amz, apz, diag, fl = H.coefficients(l)
    rhs = H.residual(l, e, np.zeros(shp)) * -1.0       # rhs = A e on fluid cells
    out = H.line_pass(l, e, rhs, 1.3)
    assert np.allclose(out, e, rtol=1e-10, atol=1e-10 * np.abs(e).max())
    e2 = rng.standard_normal(shp) * (~H.blank[l] | gh)
    r0 = np.linalg.norm(H.residual(l, e2, rhs))
    r1 = np.linalg.norm(H.residual(l, H.line_pass(l, e2, rhs, 1.0), rhs))
    assert r1 < 0.5 * r0
