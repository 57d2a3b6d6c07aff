"""ctypes wrapper of the oracle (oracle/_build/liboracle.so): the CPU
restatement of the reference pressure path used as the checker."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class MGOpts(C.Structure):
    _fields_ = [("iter_max", C.c_int), ("restol", C.c_double), ("rel_tol", C.c_double),
                ("iter_finest", C.c_int), ("iter_coarsest", C.c_int), ("omega", C.c_double),
                ("itermax_ad", C.c_int), ("restol_ad", C.c_double), ("mean_remove", C.c_int),
                ("zebra", C.c_int)]


class MGStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("linf", C.c_double), ("l2", C.c_double),
                ("b_l2", C.c_double), ("loc", C.c_int32 * 3), ("levels", C.c_int * 3)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run make -C oracle")
        L = C.CDLL(ORACLE_SO)
        P = C.c_void_p
        L.or_create.restype = P
        L.or_create.argtypes = [C.c_int] * 3 + [P] * 3
        L.or_destroy.argtypes = [P]
        L.or_field_size.restype = C.c_long
        L.or_field_size.argtypes = [P]
        L.or_get_metrics.argtypes = [P, C.c_int, P]
        L.or_set_iblank.argtypes = [P, P]
        L.or_nghost.argtypes = [P]
        L.or_get_flags.argtypes = [P] * 8
        L.or_get_ghost_list.argtypes = [P] * 4
        L.or_set_ghost_stencils.argtypes = [P] * 6
        L.or_set_fresh.argtypes = [P, P]
        L.or_calc_residual.restype = C.c_double
        L.or_calc_residual.argtypes = [P, P, P, P, P, P]
        L.or_point_jacobi.argtypes = [P, P, P, C.c_double]
        L.or_mean_remove.restype = C.c_double
        L.or_mean_remove.argtypes = [P, P]
        L.or_outer_ghost.argtypes = [P, P]
        L.or_ghost_rows.argtypes = [P, P, P, C.c_int, C.c_double]
        L.or_divergence.restype = C.c_double
        L.or_divergence.argtypes = [P, P, P, P, C.c_double, P]
        L.or_tdma.argtypes = [P] * 5 + [C.c_int, C.c_int]
        L.or_mg_solve.argtypes = [P, P, P, C.POINTER(MGOpts), C.POINTER(MGStats)]
        L.or_mg_cycle.argtypes = [P, P, P, C.POINTER(MGOpts)]
        L.or_threads.restype = C.c_int
        L.or_apply_A.argtypes = [P, P, P]
        L.or_apply_AT.argtypes = [P, P, P]
        L.or_geom_ellipsoid.argtypes = [P, C.c_int] + [C.c_double] * 7
        L.or_geom_clear_bodies.argtypes = [P]
        L.or_geom_nbodies.argtypes = [P]
        L.or_geom_surface_sizes.argtypes = [P, C.c_int, P, P]
        L.or_geom_get_surface.argtypes = [P, C.c_int, P, P, P, P]
        L.or_geom_set_boundary.argtypes = [P, P, C.c_double]
        L.or_geom_get_iblank.argtypes = [P, P, P]
        L.or_geom_get_ghost_geometry.argtypes = [P] * 10
        L.or_geom_get_dirichlet.argtypes = [P, P]
        L.or_set_vel_ghosts.argtypes = [P, P, C.c_double, C.c_double]
        L.or_vel_ghosts.argtypes = [P] * 5 + [C.c_int, C.c_double]
        L.or_correct_vel.argtypes = [P, C.c_double] + [P] * 11 + [C.c_int, C.c_double]
        L.or_update_pressure.argtypes = [P, C.c_int, P, P, P, C.c_int, C.c_double]
        L.or_face_vel.argtypes = [P, C.c_int, C.c_double] + [P] * 10
        L.or_solve_ad_cn2.restype = C.c_int
        L.or_solve_ad_cn2.argtypes = [P, C.c_double] + [P] * 11 + [C.c_int] + [P] * 4 + [C.c_int, C.c_double, P]
        L.or_geom_set_freshcell.restype = C.c_int
        L.or_geom_set_freshcell.argtypes = [P, C.c_double]
        L.or_nfresh_tables.restype = C.c_int
        L.or_nfresh_tables.argtypes = [P]
        L.or_geom_get_fresh.argtypes = [P] * 11
        L.or_set_fresh_tables.argtypes = [P, C.c_int] + [P] * 7
        L.or_set_solve_ad.argtypes = [P, C.c_double] + [P] * 5
        L.or_calc_residual_ad.restype = C.c_double
        L.or_calc_residual_ad.argtypes = [P, C.c_double] + [P] * 8
        L.or_itsolv_ad.argtypes = [P, C.c_double] + [P] * 6
        L.or_solve_ad.restype = C.c_int
        L.or_solve_ad.argtypes = [P, C.c_double] + [P] * 11 + [C.c_int, C.c_double, P]
        _L = L
    return _L


def mg_opts(iter_max=200, restol=0.0, rel_tol=1e-12, iter_finest=2, iter_coarsest=4, omega=1.0,
            itermax_ad=500, restol_ad=1e-15, mean_remove=1, zebra=0):
    return MGOpts(iter_max, restol, rel_tol, iter_finest, iter_coarsest, omega, itermax_ad,
                  restol_ad, mean_remove, zebra)


class Oracle:
    def __init__(self, xn, yn, zn):
        self.xn, self.yn, self.zn = (np.ascontiguousarray(a, dtype=np.float64) for a in (xn, yn, zn))
        self.nxc, self.nyc, self.nzc = len(xn) - 1, len(yn) - 1, len(zn) - 1
        self.h = lib().or_create(self.nxc, self.nyc, self.nzc, self.xn.ctypes.data,
                                 self.yn.ctypes.data, self.zn.ctypes.data)
        assert self.h
        self.shape = (self.nzc + 3, self.nyc + 3, self.nxc + 3)

    def __del__(self):
        try:
            lib().or_destroy(self.h)
        except Exception:
            pass

    def metrics(self, d):
        n = (self.nxc, self.nyc, self.nzc)[d] + 1
        out = np.zeros(6 * (n + 2))
        lib().or_get_metrics(self.h, d, out.ctypes.data)
        return out.reshape(6, n + 2)  # x, xc, dx, dxinv, dxc, dxcinv

    def set_iblank(self, ib):
        ib = np.ascontiguousarray(ib, dtype=np.int8)
        return lib().or_set_iblank(self.h, ib.ctypes.data)

    def flags(self):
        arrs = [np.zeros(self.shape, np.int8) for _ in range(7)]
        lib().or_get_flags(self.h, *[a.ctypes.data for a in arrs])
        return dict(zip(("ghost", "ium", "iup", "jum", "jup", "kum", "kup"), arrs))

    def ghost_list(self):
        n = lib().or_nghost(self.h)
        a = [np.zeros(max(n, 1), np.int32) for _ in range(3)]
        lib().or_get_ghost_list(self.h, *[x.ctypes.data for x in a])
        return [x[:n] for x in a]

    def set_stencils(self, gs):
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (gs.ic, gs.jc, gs.kc)]
        co = np.ascontiguousarray(gs.coeff, dtype=np.float64)
        pr = np.ascontiguousarray(gs.probe, dtype=np.float64)
        lib().or_set_ghost_stencils(self.h, *[a.ctypes.data for a in arrs], co.ctypes.data, pr.ctypes.data)
        self._keep = (arrs, co, pr)

    # ---- triangulated-body geometry (or_geom.c) ----
    def add_ellipsoid(self, center, semi, alpha_deg=0.0, n_theta=24):
        m = lib().or_geom_ellipsoid(self.h, int(n_theta), *[float(v) for v in center],
                                    *[float(v) for v in semi], float(alpha_deg))
        if m < 0:
            raise ValueError(f"or_geom_ellipsoid rejected the body ({m})")
        return m

    def clear_bodies(self):
        lib().or_geom_clear_bodies(self.h)

    def surface(self, body):
        nm, ne = C.c_int(), C.c_int()
        lib().or_geom_surface_sizes(self.h, body, C.byref(nm), C.byref(ne))
        xyz = np.zeros((nm.value, 3))
        tri = np.zeros((ne.value, 3), np.int32)
        cent = np.zeros((ne.value, 3))
        nrm = np.zeros((ne.value, 3))
        lib().or_geom_get_surface(self.h, body, xyz.ctypes.data, tri.ctypes.data, cent.ctypes.data,
                                  nrm.ctypes.data)
        return xyz, tri, cent, nrm

    def set_boundary(self, iblank_prev=None, probe_norm=2.0):
        """set_internal_boundary; returns (rc, iblank, fresh, ghosts dict)"""
        prev = None if iblank_prev is None else np.ascontiguousarray(iblank_prev, dtype=np.int8)
        rc = lib().or_geom_set_boundary(self.h, None if prev is None else prev.ctypes.data, float(probe_norm))
        ib = np.zeros(self.shape, np.int8)
        fr = np.zeros(self.shape, np.int8)
        lib().or_geom_get_iblank(self.h, ib.ctypes.data, fr.ctypes.data)
        if rc:
            return rc, ib, fr, None
        n = lib().or_nghost(self.h)
        m = max(n, 1)
        ig, jg, kg = self.ghost_list()
        g = {"ig": ig, "jg": jg, "kg": kg, "ic": np.zeros(m, np.int32), "jc": np.zeros(m, np.int32),
             "kc": np.zeros(m, np.int32), "coeff": np.zeros((m, 8)), "probe": np.zeros(m),
             "bi": np.zeros((m, 3)), "ip": np.zeros((m, 3)), "elem": np.zeros(m, np.int32),
             "body": np.zeros(m, np.int32)}
        lib().or_geom_get_ghost_geometry(self.h, *[g[k].ctypes.data for k in
                                                   ("ic", "jc", "kc", "coeff", "probe", "bi", "ip", "elem", "body")])
        g["coeffD"] = np.zeros((m, 8))
        lib().or_geom_get_dirichlet(self.h, g["coeffD"].ctypes.data)
        for k in ("ic", "jc", "kc", "coeff", "probe", "bi", "ip", "elem", "body", "coeffD"):
            g[k] = g[k][:n]
        return rc, ib, fr, g

    def set_vel_ghosts(self, coeffD, bi_weight=2.0, ip_weight=1.0):
        cd = np.ascontiguousarray(coeffD, dtype=np.float64)
        lib().or_set_vel_ghosts(self.h, cd.ctypes.data, float(bi_weight), float(ip_weight))

    def vel_ghosts(self, u, v, w, vbi=None, itermax=1000, restol=0.0):
        vb = None if vbi is None else np.ascontiguousarray(vbi, dtype=np.float64)
        return lib().or_vel_ghosts(self.h, u.ctypes.data, v.ctypes.data, w.ctypes.data,
                                   None if vb is None else vb.ctypes.data, itermax, restol)

    def correct_vel(self, dt, p, u, v, w, fu, fv, fw, bcxu, bcyv, bczw, vbi=None, itermax_ad=1000, restol_ad=0.0):
        """correct_vel on copies; returns (u, v, w, fu, fv, fw)"""
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (u, v, w, fu, fv, fw)]
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (p, bcxu, bcyv, bczw)]
        vb = None if vbi is None else np.ascontiguousarray(vbi, dtype=np.float64)
        lib().or_correct_vel(self.h, float(dt), ins[0].ctypes.data, *[a.ctypes.data for a in out],
                             *[a.ctypes.data for a in ins[1:]], None if vb is None else vb.ctypes.data,
                             itermax_ad, restol_ad)
        return out

    def update_pressure(self, van_kan, pprime, p, g=None, itermax=1000, restol=0.0):
        out = np.array(p, dtype=np.float64, order="C", copy=True)
        pp_ = np.ascontiguousarray(pprime, dtype=np.float64)
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        lib().or_update_pressure(self.h, int(van_kan), pp_.ctypes.data, out.ctypes.data,
                                 None if gg is None else gg.ctypes.data, itermax, restol)
        return out

    def face_vel(self, van_kan, dt, p, u, v, w, bcxu, bcyv, bczw, faces=None):
        """face_vel; returns (face_u, face_v, face_w), written over copies of
        `faces` (zeros when None)"""
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (p, u, v, w, bcxu, bcyv, bczw)]
        out = [np.zeros(self.shape) if faces is None else np.array(f, dtype=np.float64, order="C", copy=True)
               for f in (faces or (None, None, None))]
        lib().or_face_vel(self.h, int(van_kan), float(dt), *[a.ctypes.data for a in ins], *[a.ctypes.data for a in out])
        return out

    # ---- implicit momentum solve (or_ad.c) ----
    def set_solve_ad(self, dt, visc, bxv, byv, bzv):
        """amx_ad, apx_ad, amy_ad, apy_ad, amz_ad, apz_ad as a (6, ...) array"""
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (visc, bxv, byv, bzv)]
        coef = np.zeros((6,) + self.shape)
        lib().or_set_solve_ad(self.h, float(dt), *[a.ctypes.data for a in ins], coef.ctypes.data)
        return coef

    @staticmethod
    def _faces(faces):
        return [np.ascontiguousarray(a, dtype=np.float64) for a in faces]

    def calc_residual_ad(self, half_dt, coef, faces, var, r):
        """(max |res|, 1-based loc, res field)"""
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        f = self._faces(faces)
        var = np.ascontiguousarray(var, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        res = np.zeros(self.shape)
        loc = np.zeros(3, np.int32)
        m = lib().or_calc_residual_ad(self.h, float(half_dt), coef.ctypes.data, *[a.ctypes.data for a in f],
                                      var.ctypes.data, r.ctypes.data, res.ctypes.data, loc.ctypes.data)
        return m, tuple(int(v) for v in loc), res

    def itsolv_ad(self, half_dt, coef, faces, var, r):
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        f = self._faces(faces)
        out = np.array(var, dtype=np.float64, order="C", copy=True)
        r = np.ascontiguousarray(r, dtype=np.float64)
        lib().or_itsolv_ad(self.h, float(half_dt), coef.ctypes.data, *[a.ctypes.data for a in f], out.ctypes.data,
                           r.ctypes.data)
        return out

    def solve_ad(self, half_dt, coef, faces, u, v, w, nlu, nlv, nlw, vbi=None, itermax=1000, restol=1e-12):
        """GCM loop of solve_ad on copies; returns (u, v, w, iterations, maxres)"""
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        f = self._faces(faces)
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (u, v, w)]
        rs = [np.ascontiguousarray(a, dtype=np.float64) for a in (nlu, nlv, nlw)]
        vb = None if vbi is None else np.ascontiguousarray(vbi, dtype=np.float64)
        mr = C.c_double()
        it = lib().or_solve_ad(self.h, float(half_dt), coef.ctypes.data, *[a.ctypes.data for a in f],
                               *[a.ctypes.data for a in out], *[a.ctypes.data for a in rs],
                               None if vb is None else vb.ctypes.data, int(itermax), float(restol),
                               C.addressof(mr))
        return out[0], out[1], out[2], it, mr.value

    def solve_ad_cn2(self, dt, coef, faces, u, v, w, nlu, nlv, nlw, van_kan, p, bcxu, bcyv, bczw, vbi=None,
                     itermax=1000, restol=1e-12):
        """solve_ad with CRANK_NICOLSON2 on copies; returns (u, v, w, faces, iterations, maxres)"""
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        f = [np.array(a, dtype=np.float64, order="C", copy=True) for a in faces]
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (u, v, w)]
        rs = [np.ascontiguousarray(a, dtype=np.float64) for a in (nlu, nlv, nlw)]
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (p, bcxu, bcyv, bczw)]
        vb = None if vbi is None else np.ascontiguousarray(vbi, dtype=np.float64)
        mr = C.c_double()
        it = lib().or_solve_ad_cn2(self.h, float(dt), coef.ctypes.data, *[a.ctypes.data for a in f],
                                   *[a.ctypes.data for a in out], *[a.ctypes.data for a in rs],
                                   None if vb is None else vb.ctypes.data, int(van_kan),
                                   *[a.ctypes.data for a in ins], int(itermax), float(restol), C.addressof(mr))
        return out[0], out[1], out[2], f, it, mr.value

    def set_fresh(self, fr):
        fr = np.ascontiguousarray(fr, dtype=np.int8)
        lib().or_set_fresh(self.h, fr.ctypes.data)

    FRESH_KEYS = ("i", "j", "k", "ic", "jc", "kc", "coeffD", "bi", "body", "elem")

    def set_freshcell(self, probe_norm=2.0):
        """GCM_set_freshcell for the current fresh mask; returns (rc, tables)"""
        rc = lib().or_geom_set_freshcell(self.h, float(probe_norm))
        return rc, (self.fresh_tables() if rc == 0 else None)

    def fresh_tables(self):
        n = lib().or_nfresh_tables(self.h)
        m = max(n, 1)
        t = {k: np.zeros(m, np.int32) for k in ("i", "j", "k", "ic", "jc", "kc", "body", "elem")}
        t["coeffD"] = np.zeros((m, 8))
        t["bi"] = np.zeros((m, 3))
        lib().or_geom_get_fresh(self.h, *[t[k].ctypes.data for k in self.FRESH_KEYS])
        return {k: v[:n] for k, v in t.items()}

    def set_fresh_tables(self, t):
        a = {k: np.ascontiguousarray(t[k], dtype=np.int32) for k in ("i", "j", "k", "ic", "jc", "kc")}
        cd = np.ascontiguousarray(t["coeffD"], dtype=np.float64)
        lib().or_set_fresh_tables(self.h, len(a["i"]), *[a[k].ctypes.data for k in ("i", "j", "k", "ic", "jc", "kc")],
                                  cd.ctypes.data)

    def calc_residual(self, var, r):
        var = np.ascontiguousarray(var, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        res = np.zeros(self.shape)
        loc = np.zeros(3, np.int32)
        l2 = C.c_double()
        m = lib().or_calc_residual(self.h, var.ctypes.data, r.ctypes.data, res.ctypes.data,
                                   loc.ctypes.data, C.addressof(l2))
        return res, m, tuple(int(v) for v in loc), l2.value

    def mean_remove(self, p):
        return lib().or_mean_remove(self.h, p.ctypes.data)

    def point_jacobi(self, var, r, omega=1.0, sweeps=1):
        """sweeps of MG_itsolv_Point_Jacobi on a copy of var"""
        v = np.array(var, dtype=np.float64, order="C", copy=True)
        r = np.ascontiguousarray(r, dtype=np.float64)
        for _ in range(sweeps):
            lib().or_point_jacobi(self.h, v.ctypes.data, r.ctypes.data, float(omega))
        return v

    def outer_ghost(self, p):
        lib().or_outer_ghost(self.h, p.ctypes.data)

    def ghost_rows(self, p, g=None, itermax=1000, restol=1e-15):
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        return lib().or_ghost_rows(self.h, p.ctypes.data, None if gg is None else gg.ctypes.data,
                                   itermax, restol)

    def divergence(self, fu, fv, fw, dtinv=1.0):
        out = np.zeros(self.shape)
        s = lib().or_divergence(self.h, fu.ctypes.data, fv.ctypes.data, fw.ctypes.data, dtinv, out.ctypes.data)
        return out, s

    def apply_A(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        y = np.zeros(self.shape)
        lib().or_apply_A(self.h, p.ctypes.data, y.ctypes.data)
        return y

    def apply_AT(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.shape)
        lib().or_apply_AT(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def mg_solve(self, p, rhs, **kw):
        o = mg_opts(**kw)
        s = MGStats()
        rc = lib().or_mg_solve(self.h, p.ctypes.data, np.ascontiguousarray(rhs).ctypes.data,
                               C.byref(o), C.byref(s))
        return rc, {"iterations": s.iterations, "linf": s.linf, "l2": s.l2, "b_l2": s.b_l2,
                    "loc": tuple(s.loc), "levels": tuple(s.levels)}

    def mg_cycle(self, p, rhs, **kw):
        o = mg_opts(**kw)
        lib().or_mg_cycle(self.h, p.ctypes.data, np.ascontiguousarray(rhs).ctypes.data, C.byref(o))


def tdma(a, b, c, r):
    """Run the oracle's tdma on one system given as 0-based arrays of length n."""
    n = len(b)
    A = np.zeros(n + 1); B = np.ones(n + 1); Cc = np.zeros(n + 1); R = np.zeros(n + 1); U = np.zeros(n + 1)
    A[1:] = a; B[1:] = b; Cc[1:] = c; R[1:] = r
    lib().or_tdma(A.ctypes.data, B.ctypes.data, Cc.ctypes.data, R.ctypes.data, U.ctypes.data, 1, n)
    return U[1:]


def threads():
    """OpenMP threads the oracle's zebra variant runs on."""
    return int(lib().or_threads())


def build_flags():
    """The C compiler flags the oracle is built with (oracle/Makefile)."""
    import re
    txt = open(os.path.join(ROOT, "oracle", "Makefile")).read()
    m = re.search(r"^CFLAGS\s*\?=\s*(.*)$", txt, re.M)
    return m.group(1).strip() if m else ""
