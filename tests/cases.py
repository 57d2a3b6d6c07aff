"""Shared test cases: grids, bodies and RHS fields in the reference layout."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from flow_structure_interaction_fem_ibm_amd import pp


def uniform(n, length=1.0):
    return np.linspace(0.0, length, n + 1)


def sphere_case(n=32, r=0.15, c=(0.5, 0.5, 0.5), stretch=0.0, mode=pp.PP_GEOM_ANALYTIC, ny=None, nz=None):
    """An n^3 (or n x ny x nz) grid on [0,1]^3 with one sphere."""
    ny = ny or n
    nz = nz or n
    xn = uniform(n) if stretch == 0 else pp.tanh_nodes(n, 1.0, stretch)
    yn = uniform(ny) if stretch == 0 else pp.tanh_nodes(ny, 1.0, stretch)
    zn = uniform(nz) if stretch == 0 else pp.tanh_nodes(nz, 1.0, stretch)
    ib, fr, gs, ex = pp.build_geometry(xn, yn, zn, [pp.Ellipsoid(c, (r, r, r), 0.0, 48)], mode=mode)
    return xn, yn, zn, ib, gs, ex


def interior(a):
    return a[1:-2, 1:-2, 1:-2]


def random_field(shape, seed, ib=None):
    rng = np.random.default_rng(seed)
    f = np.zeros(shape)
    f[1:-2, 1:-2, 1:-2] = rng.uniform(-1, 1, (shape[0] - 3, shape[1] - 3, shape[2] - 3))
    if ib is not None:
        f = f * (1 - ib)
    return f


def metrics(nodes):
    """make_grid arrays (0:n+1) for one direction from nodes x(1..n)."""
    n = len(nodes)
    x = np.zeros(n + 2)
    x[1:n + 1] = nodes
    x[0] = -x[2]
    x[n + 1] = x[n] + x[n] - x[n - 1]
    xc = np.zeros(n + 2)
    dx = np.zeros(n + 2)
    xc[:n + 1] = 0.5 * (x[1:n + 2] + x[:n + 1])
    dx[:n + 1] = x[1:n + 2] - x[:n + 1]
    xc[n + 1] = x[n + 1]
    dxc = np.zeros(n + 2)
    dxc[1:n + 1] = xc[1:n + 1] - xc[:n]
    return x, xc, dx, dxc


def assemble(xn, yn, zn, ib, gs):
    """Sparse matrix of the pressure operator over the reference-layout index
    space, assembled from its definition (fluid rows: 7-point stretched
    Laplacian with Neumann outer faces; ghost rows: p_g - sum c_r p_r; solid
    rows: identity).  Independent of the oracle's and the library's code."""
    nxc, nyc, nzc = len(xn) - 1, len(yn) - 1, len(zn) - 1
    ex, ey = nxc + 3, nyc + 3
    N = ex * ey * (nzc + 3)
    mx = [metrics(a) for a in (xn, yn, zn)]
    ibf = ib.reshape(-1)
    rows, cols, vals = [], [], []
    ghost = np.zeros(N, bool)
    gidx = (gs.ig + ex * (gs.jg + ey * gs.kg)).astype(np.int64)
    ghost[gidx] = True
    K, J, I = np.meshgrid(np.arange(1, nzc + 1), np.arange(1, nyc + 1), np.arange(1, nxc + 1), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    idx = I + ex * (J + ey * K)
    fl = ibf[idx] == 0
    for d, (C, stride, n) in enumerate(zip((I, J, K), (1, ex, ex * ey), (nxc, nyc, nzc))):
        x, xc, dx, dxc = mx[d]
        am = np.where(C > 1, 1.0 / dxc[C] / dx[C], 0.0)
        ap = np.where(C < n, 1.0 / dxc[np.minimum(C + 1, n)] / dx[C], 0.0)
        rows += [idx[fl], idx[fl], idx[fl], idx[fl]]
        cols += [idx[fl] - stride, idx[fl], idx[fl] + stride, idx[fl]]
        vals += [am[fl], -am[fl], ap[fl], -ap[fl]]
    inc = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]])
    for q in range(gs.n):
        g = gidx[q]
        rows.append(np.array([g])); cols.append(np.array([g])); vals.append(np.array([1.0]))
        for r in range(8):
            ii, jj, kk = gs.ic[q] + inc[r, 0], gs.jc[q] + inc[r, 1], gs.kc[q] + inc[r, 2]
            if (ii, jj, kk) == (gs.ig[q], gs.jg[q], gs.kg[q]):
                continue
            ii, jj, kk = min(max(ii, 1), nxc), min(max(jj, 1), nyc), min(max(kk, 1), nzc)
            t = ii + ex * (jj + ey * kk)
            if ibf[t] == 1 and not ghost[t]:
                continue
            rows.append(np.array([g])); cols.append(np.array([t])); vals.append(np.array([-gs.coeff[q, r]]))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    unknown = np.zeros(N, bool)
    unknown[idx[fl]] = True
    unknown[gidx] = True
    return A, unknown


def volumes(xn, yn, zn):
    vx, vy, vz = (np.diff(a) for a in (xn, yn, zn))
    V = np.zeros((len(zn) + 2, len(yn) + 2, len(xn) + 2))
    V[1:-2, 1:-2, 1:-2] = vz[:, None, None] * vy[None, :, None] * vx[None, None, :]
    return V
