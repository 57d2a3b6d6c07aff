"""A numpy restatement of the multigrid hierarchy of the library (the level
sizes, the coarse iblank and face flags, the level metrics) and of its
single-level operators (residual, restriction, prolongation + correction,
one zebra line pass set), written from their definitions:

  * hierarchy: a direction is halved (n -> n // 2 cells) while its cell
    count is at least 4, until a level has at most 4096 cells or no
    direction halves; coarse nodes are every other fine node (MG_initial,
    CORE_MG.F90:186-205) and the last fine node, so for an odd count the
    last coarse cell spans three fine cells;
  * coarse iblank: a coarse cell is blank when all its children are (the
    3-D form of MG_Prepare_Iblank's product of the two children,
    CORE_MG.F90:732-786);
  * face flags: level 0 closes the outer faces only (BOUNDARY_SET.F90:54-92;
    under GCM the internal faces stay open); a coarse level closes the outer
    faces and, on a fluid cell, a face whose neighbour is blank:
    INT((iblank_nb - iblank)/2 + 0.5) (MG_Prepare, CORE_MG.F90:1736-1781);
  * metrics: x(0) = -x(2), x(n+2) = 2x(n+1) - x(n), centres, widths and the
    face coefficients dxcinv(face)*dxinv(cell) of make_grid
    (CORE_GRID.F90:115-149);
  * residual: rhs - A e on fluid cells, 0 on blank ones (MG_residual,
    CORE_MG.F90:2274-2387);
  * restriction: the fluid-volume-weighted mean of the children, 0 where no
    child is fluid (the masked averaging of MG_Restrict, CORE_MG.F90:2511-2538,
    over all halved directions at once);
  * prolongation: trilinear between the parent's centre and the neighbour
    centre on the child's side of it (a missing neighbour, or a child
    centred on its parent, gives the parent weight 1), renormalised over
    fluid coarse cells, added on fluid fine cells (CORE_MG.F90:3004-3022,
    1354-1355);
  * line pass: zebra line relaxation in x, y, z (colours by the parity of a
    line's two transverse indices) of the MG_itsolv rows (CORE_MG.F90:
    1945-2009), solved by Thomas elimination.

Arrays are dense over a level's cells, numpy shape (cz, cy, cx).  Test
infrastructure only; it does not import the library."""
from __future__ import annotations

import numpy as np

COARSE_MAX = 4096
MAX_LEVELS = 12
FL_BLANK, FL_GHOST, FL_IM, FL_IP, FL_JM, FL_JP, FL_KM, FL_KP = 1, 2, 4, 8, 16, 32, 64, 128


def dir_metrics(nodes):
    """Widths, centres and face coefficients (cm, cp) of the cells between
    consecutive nodes, with make_grid's mirrored end nodes."""
    X = np.asarray(nodes, dtype=np.float64)
    n = len(X) - 1
    x = np.zeros(n + 3)
    x[1:n + 2] = X
    x[0] = -x[2]
    x[n + 2] = x[n + 1] + x[n + 1] - x[n]
    xc = 0.5 * (x[1:] + x[:-1])            # centres of cells 0..n+1
    dx = x[1:] - x[:-1]
    dxinv = 1.0 / dx
    dxcinv = np.zeros(n + 2)
    dxcinv[1:] = 1.0 / (xc[1:] - xc[:-1])
    idx = np.arange(1, n + 1)
    return {"w": dx[idx], "xc": xc[idx], "cm": dxcinv[idx] * dxinv[idx], "cp": dxcinv[idx + 1] * dxinv[idx]}


class Hierarchy:
    """Levels of a grid (node arrays xn, yn, zn) and an iblank/ghost mask in
    the reference layout (nzc+3, nyc+3, nxc+3)."""

    def __init__(self, xn, yn, zn, iblank, ghost=None):
        nodes = [np.asarray(a, dtype=np.float64) for a in (xn, yn, zn)]
        dims = [len(a) - 1 for a in nodes]
        self.nodes = [nodes]
        self.dims = [tuple(dims)]
        self.cf = [(1, 1, 1)]
        while len(self.dims) < MAX_LEVELS:
            n = self.dims[-1]
            if n[0] * n[1] * n[2] <= COARSE_MAX:
                break
            cf = tuple(2 if v >= 4 else 1 for v in n)
            if cf == (1, 1, 1):
                break
            self.dims.append(tuple(v // f for v, f in zip(n, cf)))
            self.nodes.append([np.concatenate([a[:len(a) - 1:2][:(len(a) - 1) // 2], a[-1:]]) if f == 2 else a
                               for a, f in zip(self.nodes[-1], cf)])
            self.cf.append(cf)
        self.metrics = [[dir_metrics(a) for a in nd] for nd in self.nodes]
        ib = np.asarray(iblank)[1:-2, 1:-2, 1:-2].astype(bool)
        gh = np.zeros_like(ib) if ghost is None else np.asarray(ghost)[1:-2, 1:-2, 1:-2].astype(bool)
        self.blank = [ib]
        for l in range(1, len(self.dims)):
            b = self.blank[-1].astype(np.int64)
            for ax in range(3):
                b = np.minimum.reduceat(b, self.starts(l, ax), axis=ax)
            self.blank.append(b.astype(bool))
        self.flags = []
        for l, b in enumerate(self.blank):
            ibi = b.astype(np.int64)
            f = ibi * FL_BLANK + (gh.astype(np.int64) * FL_GHOST if l == 0 else 0)
            closed = [np.zeros(b.shape, bool) for _ in range(6)]
            for ax, (lo, hi) in zip((2, 1, 0), ((0, 1), (2, 3), (4, 5))):
                sl = [slice(None)] * 3
                sl[ax] = 0
                closed[lo][tuple(sl)] = True
                sl[ax] = -1
                closed[hi][tuple(sl)] = True
                if l > 0:
                    # INT((iblank_nb - iblank)/2 + 0.5): 1 only for a fluid cell with a blank neighbour
                    nb_m = np.roll(ibi, 1, axis=ax)
                    nb_p = np.roll(ibi, -1, axis=ax)
                    fm = np.floor((nb_m - ibi) / 2.0 + 0.5).astype(bool)
                    fp = np.floor((nb_p - ibi) / 2.0 + 0.5).astype(bool)
                    sl[ax] = 0
                    fm[tuple(sl)] = False
                    sl[ax] = -1
                    fp[tuple(sl)] = False
                    closed[lo] |= fm
                    closed[hi] |= fp
            for q, bit in enumerate((FL_IM, FL_IP, FL_JM, FL_JP, FL_KM, FL_KP)):
                f = f + closed[q].astype(np.int64) * bit
            self.flags.append(f.astype(np.uint8))

    def starts(self, l, ax):
        """First child (on level l-1) of every level-l cell along numpy axis ax;
        the last coarse cell also takes the cell an odd count leaves over."""
        d = 2 - ax
        return np.arange(self.dims[l][d]) * self.cf[l][d]

    @property
    def nlev(self):
        return len(self.dims)

    # ---- the level operator ----
    def coefficients(self, l):
        """(amx, apx, amy, apy, amz, apz, diag, fluid) of level l: the face
        coefficients with closed faces zeroed, diag = -(sum), in the
        reference's (acx + acy + acz) order."""
        f = self.flags[l]
        m = self.metrics[l]
        shp = f.shape
        amx = np.where(f & FL_IM, 0.0, np.broadcast_to(m[0]["cm"][None, None, :], shp))
        apx = np.where(f & FL_IP, 0.0, np.broadcast_to(m[0]["cp"][None, None, :], shp))
        amy = np.where(f & FL_JM, 0.0, np.broadcast_to(m[1]["cm"][None, :, None], shp))
        apy = np.where(f & FL_JP, 0.0, np.broadcast_to(m[1]["cp"][None, :, None], shp))
        amz = np.where(f & FL_KM, 0.0, np.broadcast_to(m[2]["cm"][:, None, None], shp))
        apz = np.where(f & FL_KP, 0.0, np.broadcast_to(m[2]["cp"][:, None, None], shp))
        diag = (-(amx + apx) + -(amy + apy)) + -(amz + apz)
        return amx, apx, amy, apy, amz, apz, diag, (f & FL_BLANK) == 0

    @staticmethod
    def _nb(e):
        """Neighbour values (x-, x+, y-, y+, z-, z+) with zeros outside."""
        p = np.pad(e, 1)
        return (p[1:-1, 1:-1, :-2], p[1:-1, 1:-1, 2:], p[1:-1, :-2, 1:-1], p[1:-1, 2:, 1:-1],
                p[:-2, 1:-1, 1:-1], p[2:, 1:-1, 1:-1])

    def residual(self, l, e, rhs):
        amx, apx, amy, apy, amz, apz, diag, fl = self.coefficients(l)
        xm, xp, ym, yp, zm, zp = self._nb(e)
        r = rhs - e * diag - xm * amx - xp * apx - ym * amy - yp * apy - zm * amz - zp * apz
        return np.where(fl, r, 0.0)

    def volumes(self, l):
        m = self.metrics[l]
        return m[2]["w"][:, None, None] * m[1]["w"][None, :, None] * m[0]["w"][None, None, :]

    def restrict(self, l, r):
        """Fluid-volume-weighted mean of level l's field over the children of
        each level l+1 cell; 0 where no child is fluid."""
        v = np.where(self.blank[l], 0.0, self.volumes(l))
        num, den = v * r, v
        for ax in range(3):
            num = np.add.reduceat(num, self.starts(l + 1, ax), axis=ax)
            den = np.add.reduceat(den, self.starts(l + 1, ax), axis=ax)
        return np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)

    def prolong_weights(self, l, d):
        """Per fine index of direction d: parent, other coarse index, weight of
        the parent (trilinear between the two coarse centres)."""
        xf = self.metrics[l][d]["xc"]
        xcc = self.metrics[l + 1][d]["xc"]
        nf, nc = len(xf), len(xcc)
        if self.cf[l + 1][d] == 1:
            i = np.arange(nf)
            return i, i, np.ones(nf)
        par = np.minimum(np.arange(nf) // 2, nc - 1)
        oth = np.where(xf < xcc[par], par - 1, np.where(xf > xcc[par], par + 1, par))
        w = np.ones(nf)
        ok = (oth >= 0) & (oth < nc) & (oth != par)
        o = np.where(ok, oth, par)
        w[ok] = (xf[ok] - xcc[oth[ok]]) / (xcc[par[ok]] - xcc[oth[ok]])
        return par, o, w

    def prolong(self, l, ec, ef):
        """ef + interpolation of the level l+1 correction ec on fluid cells."""
        P = [self.prolong_weights(l, d) for d in range(3)]
        cfl = ~self.blank[l + 1]
        num = np.zeros(ef.shape)
        den = np.zeros(ef.shape)
        K, J, I = np.meshgrid(np.arange(ef.shape[0]), np.arange(ef.shape[1]), np.arange(ef.shape[2]), indexing="ij")
        for a in range(2):
            ci = P[0][0][I] if a == 0 else P[0][1][I]
            wi = P[0][2][I] if a == 0 else 1.0 - P[0][2][I]
            for b in range(2):
                cj = P[1][0][J] if b == 0 else P[1][1][J]
                wj = P[1][2][J] if b == 0 else 1.0 - P[1][2][J]
                for c in range(2):
                    ck = P[2][0][K] if c == 0 else P[2][1][K]
                    wk = P[2][2][K] if c == 0 else 1.0 - P[2][2][K]
                    w = wi * wj * wk
                    use = (w != 0.0) & cfl[ck, cj, ci]
                    num += np.where(use, w * ec[ck, cj, ci], 0.0)
                    den += np.where(use, w, 0.0)
        add = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
        return np.where(self.blank[l], ef, ef + add)

    def line_pass(self, l, e, rhs, omega):
        """One zebra pass set (x, y, z lines, colour 0 then 1)."""
        amx, apx, amy, apy, amz, apz, diag, fl = self.coefficients(l)
        ghost = (self.flags[l] & FL_GHOST) != 0
        blank = ~fl
        e = e.copy()
        cz, cy, cx = e.shape
        K, J, I = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
        terms = [(amx, apx), (amy, apy), (amz, apz)]
        for d, colour_of in ((0, (J + K) % 2), (1, (I + K) % 2), (2, (I + J) % 2)):
            for col in (0, 1):
                nb = self._nb(e)
                r = rhs.copy()
                for q in range(3):
                    if q != d:
                        r = r - nb[2 * q] * terms[q][0] - nb[2 * q + 1] * terms[q][1]
                a = np.where(blank, 0.0, terms[d][0])
                c = np.where(blank, 0.0, terms[d][1])
                bb = np.where(blank, 1.0, diag)
                dd = np.where(blank, np.where(ghost, e, 0.0), r)
                guard = (~blank) & (np.abs(diag) < 1e-12)
                a, c = np.where(guard, 0.0, a), np.where(guard, 0.0, c)
                bb, dd = np.where(guard, 1.0, bb), np.where(guard, e, dd)
                ax = 2 - d
                sel = np.moveaxis(colour_of == col, ax, -1)[..., 0]
                A_, B_, C_, D_ = (np.moveaxis(v, ax, -1)[sel] for v in (a, bb, c, dd))
                x = thomas(A_, B_, C_, D_)
                new = np.moveaxis(e.copy(), ax, -1)
                new[sel] = (1.0 - omega) * new[sel] + omega * x
                e = np.moveaxis(new, -1, ax)
        return e


def thomas(a, b, c, d):
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
