"""GPU: the finest-level smoothers through pp_smooth.  The point-Jacobi sweep
(MG_itsolv_Point_Jacobi) is bit-exact against the oracle; the zebra line
pass set is checked against a numpy restatement of the same line rows
(MG_itsolv) solved by Thomas elimination; a solve with the point-Jacobi
smoother reaches the line smoother's solution."""
import numpy as np
import pytest

import cases
from flow_structure_interaction_fem_ibm_amd import pp
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[case] {msg}", flush=True)


def setup(n, stretch, r=0.22, ny=None, nz=None):
    xn, yn, zn, ib, gs, _ = cases.sphere_case(n, r=r, stretch=stretch, ny=ny, nz=nz)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    o = Oracle(xn, yn, zn)
    o.set_iblank(ib)
    o.set_stencils(gs)
    return s, o, ib, gs


@pytest.mark.parametrize("n,ny,nz,sweeps,omega", [(32, None, None, 1, 1.0), (32, None, None, 3, 0.8),
                                                  (22, 20, 18, 2, 0.7)])
def test_point_jacobi_matches_oracle_bitwise(n, ny, nz, sweeps, omega):
    announce(f"point Jacobi {n}x{ny or n}x{nz or n} sweeps={sweeps} omega={omega}")
    s, o, ib, gs = setup(n, 1.5, ny=ny, nz=nz)
    p = cases.random_field(ib.shape, 11)
    b = cases.random_field(ib.shape, 12, ib)
    out = s.smooth(p, b, pp.PP_SMOOTH_POINT_JACOBI, omega, sweeps)
    ref = o.point_jacobi(p, b, omega, sweeps)
    assert np.array_equal(out, ref)


def _thomas(a, b, c, d):
    """Batched Thomas elimination over the last axis (rows are systems)."""
    n = b.shape[-1]
    cp = np.zeros_like(b)
    dp = np.zeros_like(b)
    cp[:, 0] = c[:, 0] / b[:, 0]
    dp[:, 0] = d[:, 0] / b[:, 0]
    for m in range(1, n):
        den = b[:, m] - a[:, m] * cp[:, m - 1]
        cp[:, m] = c[:, m] / den
        dp[:, m] = (d[:, m] - a[:, m] * dp[:, m - 1]) / den
    x = np.zeros_like(b)
    x[:, -1] = dp[:, -1]
    for m in range(n - 2, -1, -1):
        x[:, m] = dp[:, m] - cp[:, m] * x[:, m + 1]
    return x


def zebra_line_sweep(o, ib, ghost, var, rhs, omega):
    """One zebra pass set (x, y, z lines, colours 0 then 1; the colour of an
    x-line is (j + k) % 2, of a y-line (i + k) % 2, of a z-line (i + j) % 2)
    of the MG_itsolv line rows (CORE_MG.F90:1945-2009): fluid rows are the
    stretched 7-point Laplacian with the outer faces closed, blank rows hold
    the ghost value (0 on other blank cells), then the SOR blend."""
    e = var.copy()
    nz, ny, nx = (s - 3 for s in ib.shape)
    met = [o.metrics(d) for d in range(3)]   # rows: x, xc, dx, dxinv, dxc, dxcinv

    def coef(d, n):
        m = met[d]
        idx = np.arange(1, n + 1)
        am = m[5][idx] * m[3][idx]
        ap = m[5][idx + 1] * m[3][idx]
        am[0] = 0.0
        ap[-1] = 0.0
        return am, ap

    cmx, cpx = coef(0, nx)
    cmy, cpy = coef(1, ny)
    cmz, cpz = coef(2, nz)
    I = (slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1))
    blank = ib[I] == 1
    gh = ghost[I]
    amx = np.broadcast_to(cmx[None, None, :], blank.shape)
    apx = np.broadcast_to(cpx[None, None, :], blank.shape)
    amy = np.broadcast_to(cmy[None, :, None], blank.shape)
    apy = np.broadcast_to(cpy[None, :, None], blank.shape)
    amz = np.broadcast_to(cmz[:, None, None], blank.shape)
    apz = np.broadcast_to(cpz[:, None, None], blank.shape)
    diag = (-(amx + apx) + -(amy + apy)) + -(amz + apz)
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for d, colour_of in ((0, (jj + kk) % 2), (1, (ii + kk) % 2), (2, (ii + jj) % 2)):
        for col in (0, 1):
            ec = e[I]
            xm, xp = e[1:nz + 1, 1:ny + 1, 0:nx], e[1:nz + 1, 1:ny + 1, 2:nx + 2]
            ym, yp = e[1:nz + 1, 0:ny, 1:nx + 1], e[1:nz + 1, 2:ny + 2, 1:nx + 1]
            zm, zp = e[0:nz, 1:ny + 1, 1:nx + 1], e[2:nz + 2, 1:ny + 1, 1:nx + 1]
            terms = [(amx, xm, apx, xp), (amy, ym, apy, yp), (amz, zm, apz, zp)]
            r = rhs[I].copy()
            for q in range(3):
                if q != d:
                    r = r - terms[q][1] * terms[q][0] - terms[q][3] * terms[q][2]
            a = np.where(blank, 0.0, terms[d][0])
            c = np.where(blank, 0.0, terms[d][2])
            bb = np.where(blank, 1.0, diag)
            dd = np.where(blank, np.where(gh, ec, 0.0), r)
            # move the line axis last and keep the lines of this colour
            ax = 2 - d
            sel = np.moveaxis(colour_of == col, ax, -1)[..., 0]
            A_, B_, C_, D_ = (np.moveaxis(v, ax, -1)[sel] for v in (a, bb, c, dd))
            x = _thomas(A_, B_, C_, D_)
            new = np.moveaxis(ec.copy(), ax, -1)
            new[sel] = (1.0 - omega) * new[sel] + omega * x
            e[I] = np.moveaxis(new, -1, ax)
    return e


@pytest.mark.parametrize("n,ny,nz,omega", [(32, None, None, 1.0), (24, 20, 22, 0.9)])
def test_line_pass_set_matches_numpy_restatement(n, ny, nz, omega):
    announce(f"zebra line pass set {n}x{ny or n}x{nz or n} omega={omega}")
    s, o, ib, gs = setup(n, 1.5, ny=ny, nz=nz)
    p = cases.random_field(ib.shape, 21)
    b = cases.random_field(ib.shape, 22, ib)
    ghost = np.zeros(ib.shape, bool)
    ghost[gs.kg, gs.jg, gs.ig] = True
    out = s.smooth(p, b, pp.PP_SMOOTH_LINE, omega, 1)
    ref = zebra_line_sweep(o, ib, ghost, p, b, omega)
    scale = np.abs(ref).max()
    assert np.allclose(out, ref, rtol=1e-11, atol=1e-12 * scale)
    # two sweeps = the restatement applied twice
    out2 = s.smooth(p, b, pp.PP_SMOOTH_LINE, omega, 2)
    ref2 = zebra_line_sweep(o, ib, ghost, ref, b, omega)
    assert np.allclose(out2, ref2, rtol=1e-11, atol=1e-12 * np.abs(ref2).max())


@pytest.mark.parametrize("prec", [pp.PP_PREC_FP64, pp.PP_PREC_MIXED])
def test_solve_with_point_jacobi_smoother(prec):
    announce(f"solve with point Jacobi smoother 32^3 prec={prec}")
    xn, yn, zn, ib, gs, _ = cases.sphere_case(32, r=0.2)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    rhs = cases.random_field(ib.shape, 31, ib)
    p_line = np.zeros(ib.shape)
    st_line = s.solve(p_line, rhs, rel_tol=1e-11, max_outer=200, precision=prec)
    p_jac = np.zeros(ib.shape)
    st_jac = s.solve(p_jac, rhs, rel_tol=1e-11, max_outer=200, precision=prec,
                     smoother=pp.PP_SMOOTH_POINT_JACOBI, omega=0.8)
    assert st_line["status"] == "PP_OK" and st_jac["status"] == "PP_OK"
    assert st_jac["outer_iterations"] >= st_line["outer_iterations"]
    fl = np.zeros(ib.shape, bool)
    fl[1:-2, 1:-2, 1:-2] = ib[1:-2, 1:-2, 1:-2] == 0
    err = np.linalg.norm((p_jac - p_line)[fl]) / np.linalg.norm(p_line[fl])
    assert err < 1e-8, err
    print(f"[result] iterations line {st_line['outer_iterations']} jacobi {st_jac['outer_iterations']} "
          f"diff {err:.2e}", flush=True)


def test_smooth_rejects_bad_arguments():
    xn, yn, zn, ib, gs, _ = cases.sphere_case(16, r=0.2)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    p = np.zeros(ib.shape)
    with pytest.raises(pp.PPError):
        s.smooth(p, p, smoother=7)
    with pytest.raises(pp.PPError):
        s.smooth(p, p, omega=2.5)
    with pytest.raises(pp.PPError):
        s.solve(p, p, smoother=3)
