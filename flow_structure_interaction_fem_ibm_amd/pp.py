"""ctypes binding of the pressure-Poisson C-ABI (include/pp_hip.h).

Mirrors the reference's pressure path (IBM/CORE_POISSON.F90 solve_poisson,
calc_residual, subtracting_average_pressure, GCM_correct_pressure;
IBM/CORE_TDMA.F90 tdma) on numpy arrays in the reference layout
(0:nx+1, 0:ny+1, 0:nz+1), i fastest, i.e. numpy shape (nzc+3, nyc+3, nxc+3).
Device-resident solves take torch tensors on a ROCm device.

The library is loaded from this package directory; nothing falls back to a
CPU path: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpp_hip.so")

PP_OK = 0
PP_NOT_CONVERGED = 1
PP_BAD_ARG = 2
PP_BAD_GEOMETRY = 3
PP_HIP_ERROR = 4
PP_NONFINITE = 5
PP_NO_GEOMETRY = 6
PP_BC_NEUMANN = 2
PP_PREC_FP64 = 0
PP_PREC_MIXED = 1

_STATUS = {0: "PP_OK", 1: "PP_NOT_CONVERGED", 2: "PP_BAD_ARG", 3: "PP_BAD_GEOMETRY",
           4: "PP_HIP_ERROR", 5: "PP_NONFINITE", 6: "PP_NO_GEOMETRY"}


class PPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("nxc", C.c_int32), ("nyc", C.c_int32), ("nzc", C.c_int32),
                ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
                ("z", C.POINTER(C.c_double)), ("bc", C.c_int32 * 6), ("device", C.c_int32)]


class Geometry(C.Structure):
    _fields_ = [("iblank", C.POINTER(C.c_int8)), ("fresh_cell", C.POINTER(C.c_int8)),
                ("n_ghost", C.c_int32),
                ("i_ghost", C.POINTER(C.c_int32)), ("j_ghost", C.POINTER(C.c_int32)),
                ("k_ghost", C.POINTER(C.c_int32)),
                ("i_cell", C.POINTER(C.c_int32)), ("j_cell", C.POINTER(C.c_int32)),
                ("k_cell", C.POINTER(C.c_int32)),
                ("coeff", C.POINTER(C.c_double)), ("probe_length", C.POINTER(C.c_double))]


class BodyDesc(C.Structure):
    _fields_ = [("shape", C.c_int32), ("center", C.c_double * 3), ("semi", C.c_double * 3),
                ("alpha_deg", C.c_double), ("n_theta", C.c_int32)]


class SurfaceDesc(C.Structure):
    _fields_ = [("n_markers", C.c_int32), ("n_elems", C.c_int32), ("xyz", C.c_void_p), ("tri", C.c_void_p),
                ("outside", C.c_double * 3)]


PP_BODY_ELLIPSOID = 3
PP_GEOM_ANALYTIC = 0
PP_GEOM_TRIANGULATED = 1


class SolveOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("rel_tol", C.c_double), ("abs_linf_tol", C.c_double),
                ("max_outer", C.c_int32), ("cycles_per_outer", C.c_int32),
                ("nu_pre", C.c_int32), ("nu_post", C.c_int32), ("omega", C.c_double),
                ("mean_remove", C.c_int32), ("project_rhs", C.c_int32), ("use_graph", C.c_int32),
                ("profile", C.c_int32), ("krylov", C.c_int32),
                ("nu_coarse", C.c_int32), ("smoother", C.c_int32)]


PP_SMOOTH_LINE = 0
PP_SMOOTH_POINT_JACOBI = 1
PP_SMOOTH_LINE_NATURAL = 2   # PP_SMOOTH_LINE without the colour-split finest-level copy
PP_LOP_RESIDUAL, PP_LOP_RESTRICT, PP_LOP_RESRESTRICT, PP_LOP_PROLONG, PP_LOP_LINE_PASS = range(5)

KERNEL_CLASSES = ("outer_residual", "update", "xline", "yline", "zline", "residual", "restrict",
                  "prolong", "ghost", "coarse", "other", "krylov", "jacobi")
NK = len(KERNEL_CLASSES)


PP_ADAMS_BASHFORTH2, PP_CRANK_NICOLSON1, PP_CRANK_NICOLSON2 = 1, 2, 3


class ADStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("max_residual", C.c_double),
                ("residual_uvw", C.c_double * 3), ("loc", C.c_int32 * 3), ("ms", C.c_double)]

    def as_dict(self):
        return {"status": _STATUS.get(self.status, self.status), "iterations": self.iterations,
                "max_residual": self.max_residual, "residual_uvw": list(self.residual_uvw),
                "loc": list(self.loc), "ms": self.ms}


class SolveStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("outer_iterations", C.c_int32), ("vcycles", C.c_int32),
                ("rel_l2", C.c_double), ("abs_linf", C.c_double), ("loc", C.c_int32 * 3),
                ("b_l2", C.c_double), ("rhs_shift", C.c_double), ("mean_removed", C.c_double),
                ("ms", C.c_double), ("bytes", C.c_double), ("history", C.c_double * 64),
                ("kern_ms", C.c_double * NK), ("kern_bytes", C.c_double * NK),
                ("kern_launches", C.c_int32 * NK), ("kern_ms_l0", C.c_double * NK),
                ("kern_bytes_l0", C.c_double * NK), ("kern_launches_l0", C.c_int32 * NK)]

    def as_dict(self):
        n = min(self.outer_iterations, 64)
        return {"status": _STATUS.get(self.status, self.status),
                "outer_iterations": self.outer_iterations, "vcycles": self.vcycles,
                "rel_l2": self.rel_l2, "abs_linf": self.abs_linf, "loc": list(self.loc),
                "b_l2": self.b_l2, "rhs_shift": self.rhs_shift,
                "mean_removed": self.mean_removed, "ms": self.ms, "bytes": self.bytes,
                "history": [self.history[i] for i in range(n)],
                "kernels": {name: {"ms": self.kern_ms[q], "bytes": self.kern_bytes[q],
                                   "launches": self.kern_launches[q], "ms_l0": self.kern_ms_l0[q],
                                   "bytes_l0": self.kern_bytes_l0[q],
                                   "launches_l0": self.kern_launches_l0[q]}
                            for q, name in enumerate(KERNEL_CLASSES)}}


_lib = None


def lib():
    """Load libpp_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.pp_last_error.restype = C.c_char_p
    L.pp_default_opts.argtypes = [C.POINTER(SolveOpts)]
    L.pp_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(GridDesc)]
    L.pp_destroy.argtypes = [P]
    L.pp_set_geometry.argtypes = [P, C.POINTER(Geometry)]
    L.pp_solve.argtypes = [P, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.pp_solve_device.argtypes = [P, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.pp_adjoint.argtypes = [P, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.pp_adjoint_device.argtypes = [P, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]
    L.pp_calc_residual.argtypes = [P, P, P, P, dp, C.POINTER(C.c_int32), dp]
    L.pp_ghost_rows.argtypes = [P, P, P]
    L.pp_mean_remove.argtypes = [P, P, dp]
    L.pp_divergence.argtypes = [P, P, P, P, C.c_double, P, dp]
    L.pp_tdma_batched.argtypes = [P, P, P, P, P, P, C.c_int32, C.c_int32, C.c_int32]
    L.pp_smooth.argtypes = [P, C.c_int32, C.c_double, C.c_int32, P, P]
    L.pp_null_vector.argtypes = [P, P]
    L.pp_level_info.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.pp_time_kernel.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp, dp]
    L.pp_setup_info.argtypes = [P, dp, C.POINTER(C.c_int32), dp]
    L.pp_level_apply.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_double, P, P, P]
    L.pp_divergence_noalloc.argtypes = [P, P, P, P, C.c_double, P, dp]
    L.pp_level_data.argtypes = [P, C.c_int32, P, P, P, P, P]
    L.pp_geom_build.argtypes = [C.POINTER(C.c_void_p), C.POINTER(GridDesc), C.c_int32,
                                C.POINTER(BodyDesc), C.c_int32, C.c_double, P]
    L.pp_geom_counts.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.pp_geom_get.argtypes = [P] * 14
    L.pp_geom_ghost_elements.argtypes = [P, P, P]
    L.pp_geom_destroy.argtypes = [P]
    L.pp_geom_dirichlet.argtypes = [P, P]
    L.pp_geom_fresh.argtypes = [P] * 11
    L.pp_set_fresh_cells.argtypes = [P, C.c_int32] + [P] * 7
    L.pp_geom_build_surfaces.argtypes = [C.POINTER(C.c_void_p), C.POINTER(GridDesc), C.c_int32,
                                         C.POINTER(SurfaceDesc), C.c_double, P]
    L.pp_geom_build_device.argtypes = [C.POINTER(C.c_void_p), C.POINTER(GridDesc), C.c_int32,
                                       C.POINTER(BodyDesc), C.c_double, P]
    L.pp_geom_build_surfaces_device.argtypes = L.pp_geom_build_surfaces.argtypes
    L.pp_geom_canonical_surface.argtypes = [C.POINTER(BodyDesc), C.POINTER(C.c_int32), C.POINTER(C.c_int32), P, P]
    L.pp_set_vel_ghosts.argtypes = [P, P, C.c_double, C.c_double]
    L.pp_correct_vel.argtypes = [P, C.c_double] + [P] * 11
    L.pp_correct_vel_device.argtypes = [P, C.c_double] + [P] * 11
    L.pp_update_pressure.argtypes = [P, C.c_int32, P, P, P]
    L.pp_update_pressure_device.argtypes = [P, C.c_int32, P, P, P]
    L.pp_set_solve_ad.argtypes = [P, C.c_double] + [P] * 4
    L.pp_set_solve_ad_device.argtypes = [P, C.c_double] + [P] * 4
    L.pp_solve_ad.argtypes = [P, C.c_int32] + [P] * 10 + [C.c_int32, C.c_double, C.POINTER(ADStats)]
    L.pp_solve_ad_device.argtypes = [P, C.c_int32] + [P] * 10 + [C.c_int32, C.c_double, C.POINTER(ADStats)]
    L.pp_calc_residual_ad.argtypes = [P, C.c_int32] + [P] * 6 + [dp, C.POINTER(C.c_int32)]
    L.pp_face_vel.argtypes = [P, C.c_int32, C.c_double] + [P] * 10
    L.pp_face_vel_device.argtypes = [P, C.c_int32, C.c_double] + [P] * 10
    L.pp_solve_ad_cn2.argtypes = [P] + [P] * 9 + [C.c_int32] + [P] * 5 + [C.c_int32, C.c_double, C.POINTER(ADStats)]
    L.pp_solve_ad_cn2_device.argtypes = L.pp_solve_ad_cn2.argtypes
    for name in ("pp_create", "pp_destroy", "pp_set_geometry", "pp_solve", "pp_solve_device",
                 "pp_adjoint", "pp_calc_residual", "pp_ghost_rows", "pp_mean_remove",
                 "pp_divergence", "pp_smooth", "pp_tdma_batched", "pp_null_vector", "pp_level_info", "pp_time_kernel",
                 "pp_setup_info", "pp_geom_build", "pp_geom_counts", "pp_geom_get", "pp_geom_ghost_elements",
                 "pp_geom_destroy", "pp_geom_dirichlet", "pp_set_vel_ghosts", "pp_correct_vel",
                 "pp_correct_vel_device", "pp_update_pressure", "pp_update_pressure_device", "pp_adjoint_device",
                 "pp_set_solve_ad", "pp_set_solve_ad_device", "pp_solve_ad", "pp_solve_ad_device",
                 "pp_calc_residual_ad", "pp_geom_build_surfaces", "pp_geom_canonical_surface",
                 "pp_level_apply", "pp_level_data", "pp_divergence_noalloc", "pp_geom_fresh", "pp_set_fresh_cells",
                 "pp_face_vel", "pp_face_vel_device", "pp_solve_ad_cn2", "pp_solve_ad_cn2_device",
                 "pp_geom_build_device", "pp_geom_build_surfaces_device"):
        getattr(L, name).restype = C.c_int32
    _lib = L
    return L


EXPORTED = ("pp_default_opts", "pp_last_error", "pp_create", "pp_destroy", "pp_set_geometry",
            "pp_solve", "pp_solve_device", "pp_adjoint", "pp_calc_residual", "pp_ghost_rows",
            "pp_mean_remove", "pp_divergence", "pp_smooth", "pp_tdma_batched", "pp_null_vector",
            "pp_level_info", "pp_time_kernel", "pp_setup_info", "pp_geom_build", "pp_geom_counts", "pp_geom_get",
            "pp_geom_ghost_elements", "pp_geom_destroy", "pp_geom_dirichlet", "pp_set_vel_ghosts",
            "pp_correct_vel", "pp_correct_vel_device", "pp_update_pressure", "pp_update_pressure_device",
            "pp_adjoint_device", "pp_set_solve_ad", "pp_set_solve_ad_device", "pp_solve_ad", "pp_solve_ad_device",
            "pp_calc_residual_ad", "pp_geom_build_surfaces", "pp_geom_canonical_surface", "pp_level_apply",
            "pp_level_data", "pp_divergence_noalloc", "pp_geom_fresh", "pp_set_fresh_cells", "pp_face_vel",
            "pp_face_vel_device", "pp_solve_ad_cn2", "pp_solve_ad_cn2_device", "pp_geom_build_device",
            "pp_geom_build_surfaces_device")


def _check(rc, allow=(PP_OK,)):
    if rc not in allow:
        raise PPError(rc, lib().pp_last_error().decode())
    return rc


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def default_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    lib().pp_default_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


@dataclass
class GhostSet:
    """The GCM ghost-cell geometry of AMODULES.F90:507-579 (1-based indices)."""
    ig: np.ndarray
    jg: np.ndarray
    kg: np.ndarray
    ic: np.ndarray
    jc: np.ndarray
    kc: np.ndarray
    coeff: np.ndarray      # (n, 8): coeffGCMN(:, n)
    probe: np.ndarray      # (n,)

    @property
    def n(self):
        return int(self.ig.shape[0])


class PressureSolver:
    """One context of the MI355X pressure solve (pp_create ... pp_destroy)."""

    def __init__(self, xn, yn, zn, device=0):
        L = lib()
        self.xn = np.ascontiguousarray(xn, dtype=np.float64)
        self.yn = np.ascontiguousarray(yn, dtype=np.float64)
        self.zn = np.ascontiguousarray(zn, dtype=np.float64)
        self.nxc, self.nyc, self.nzc = len(self.xn) - 1, len(self.yn) - 1, len(self.zn) - 1
        g = GridDesc()
        g.nxc, g.nyc, g.nzc = self.nxc, self.nyc, self.nzc
        g.x, g.y, g.z = _ptr(self.xn, C.c_double), _ptr(self.yn, C.c_double), _ptr(self.zn, C.c_double)
        for f in range(6):
            g.bc[f] = PP_BC_NEUMANN
        g.device = device
        h = C.c_void_p()
        _check(L.pp_create(C.byref(h), C.byref(g)))
        self.h = h
        self.shape = (self.nzc + 3, self.nyc + 3, self.nxc + 3)
        self.ghosts = None

    def close(self):
        if getattr(self, "h", None):
            lib().pp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_geometry(self, iblank, ghosts: GhostSet, fresh=None):
        ib = np.ascontiguousarray(iblank, dtype=np.int8)
        assert ib.shape == self.shape
        gs = ghosts
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (gs.ig, gs.jg, gs.kg, gs.ic, gs.jc, gs.kc)]
        coeff = np.ascontiguousarray(gs.coeff, dtype=np.float64).reshape(-1)
        probe = np.ascontiguousarray(gs.probe, dtype=np.float64)
        geo = Geometry()
        geo.iblank = _ptr(ib, C.c_int8)
        fr = None
        if fresh is not None:
            fr = np.ascontiguousarray(fresh, dtype=np.int8)
            geo.fresh_cell = _ptr(fr, C.c_int8)
        geo.n_ghost = gs.n
        geo.i_ghost, geo.j_ghost, geo.k_ghost = (_ptr(a, C.c_int32) for a in arrs[:3])
        geo.i_cell, geo.j_cell, geo.k_cell = (_ptr(a, C.c_int32) for a in arrs[3:])
        geo.coeff = _ptr(coeff, C.c_double)
        geo.probe_length = _ptr(probe, C.c_double)
        _check(lib().pp_set_geometry(self.h, C.byref(geo)))
        self.ghosts = gs
        self._keep = (ib, arrs, coeff, probe, fr)

    def solve(self, p, rhs, bc_src=None, **kw):
        """solve_poisson on host arrays; p is updated in place.  Returns stats."""
        o = default_opts(**kw)
        st = SolveStats()
        assert p.dtype == np.float64 and p.flags.c_contiguous and p.shape == self.shape
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        bc = None if bc_src is None else np.ascontiguousarray(bc_src, dtype=np.float64)
        rc = lib().pp_solve(self.h, p.ctypes.data, r.ctypes.data,
                            None if bc is None else bc.ctypes.data, C.byref(o), C.byref(st))
        _check(rc, (PP_OK, PP_NOT_CONVERGED))
        return st.as_dict()

    def solve_device(self, p, rhs, bc_src=None, **kw):
        """Device-resident solve on torch float64 tensors in the reference layout."""
        o = default_opts(**kw)
        st = SolveStats()
        rc = lib().pp_solve_device(self.h, p.data_ptr(), rhs.data_ptr(),
                                   None if bc_src is None else bc_src.data_ptr(),
                                   C.byref(o), C.byref(st))
        _check(rc, (PP_OK, PP_NOT_CONVERGED))
        return st.as_dict()

    def adjoint(self, w, bc_grad=False, **kw):
        o = default_opts(**kw)
        st = SolveStats()
        wv = np.ascontiguousarray(w, dtype=np.float64)
        lam = np.zeros(self.shape)
        dg = np.zeros(max(self.ghosts.n if self.ghosts else 0, 1)) if bc_grad else None
        rc = lib().pp_adjoint(self.h, wv.ctypes.data, lam.ctypes.data,
                              None if dg is None else dg.ctypes.data, C.byref(o), C.byref(st))
        _check(rc, (PP_OK, PP_NOT_CONVERGED))
        if bc_grad:
            return lam, dg[: self.ghosts.n], st.as_dict()
        return lam, st.as_dict()

    def adjoint_device(self, w, lam, dj_dg=None, **kw):
        """pp_adjoint_device on torch float64 tensors: lam (reference layout) and
        dj_dg (n_ghost, optional) are written in place; returns the stats."""
        o = default_opts(**kw)
        st = SolveStats()
        rc = lib().pp_adjoint_device(self.h, w.data_ptr(), lam.data_ptr(),
                                     None if dj_dg is None else dj_dg.data_ptr(), C.byref(o), C.byref(st))
        _check(rc, (PP_OK, PP_NOT_CONVERGED))
        return st.as_dict()

    def calc_residual(self, var, r):
        var = np.ascontiguousarray(var, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        res = np.zeros(self.shape)
        resm = C.c_double()
        l2 = C.c_double()
        loc = (C.c_int32 * 3)()
        _check(lib().pp_calc_residual(self.h, var.ctypes.data, r.ctypes.data, res.ctypes.data,
                                      C.byref(resm), loc, C.byref(l2)))
        return res, resm.value, tuple(loc), l2.value

    def ghost_rows(self, p, bc_src=None):
        assert p.dtype == np.float64 and p.flags.c_contiguous
        bc = None if bc_src is None else np.ascontiguousarray(bc_src, dtype=np.float64)
        _check(lib().pp_ghost_rows(self.h, p.ctypes.data, None if bc is None else bc.ctypes.data))
        return p

    def set_vel_ghosts(self, coeffD, bi_weight=2.0, ip_weight=1.0):
        """pp_set_vel_ghosts: coeffGCMD (n_ghost x 8) and the GCM weights."""
        cd = np.ascontiguousarray(coeffD, dtype=np.float64)
        _check(lib().pp_set_vel_ghosts(self.h, cd.ctypes.data if cd.size else None, float(bi_weight),
                                       float(ip_weight)))

    def correct_vel(self, dt, p, u, v, w, face_u, face_v, face_w, bcxu, bcyv, bczw, vel_bi=None):
        """pp_correct_vel on copies of the host fields; returns (u, v, w, face_u, face_v, face_w)."""
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (u, v, w, face_u, face_v, face_w)]
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (p, bcxu, bcyv, bczw)]
        for a in out + ins:
            assert a.shape == self.shape
        vb = None if vel_bi is None else np.ascontiguousarray(vel_bi, dtype=np.float64)
        _check(lib().pp_correct_vel(self.h, float(dt), ins[0].ctypes.data, *[a.ctypes.data for a in out],
                                    *[a.ctypes.data for a in ins[1:]], None if vb is None else vb.ctypes.data))
        return out

    def correct_vel_device(self, dt, p, u, v, w, face_u, face_v, face_w, bcxu, bcyv, bczw, vel_bi=None):
        """pp_correct_vel_device on torch float64 tensors (updated in place)."""
        t = (p, u, v, w, face_u, face_v, face_w, bcxu, bcyv, bczw)
        _check(lib().pp_correct_vel_device(self.h, float(dt), *[a.data_ptr() for a in t],
                                           None if vel_bi is None else vel_bi.data_ptr()))

    def update_pressure(self, van_kan, pprime, p, bc_src=None):
        """pp_update_pressure on a copy of p; returns the new p."""
        out = np.array(p, dtype=np.float64, order="C", copy=True)
        pr = np.ascontiguousarray(pprime, dtype=np.float64)
        bc = None if bc_src is None else np.ascontiguousarray(bc_src, dtype=np.float64)
        _check(lib().pp_update_pressure(self.h, int(van_kan), pr.ctypes.data, out.ctypes.data,
                                        None if bc is None else bc.ctypes.data))
        return out

    def update_pressure_device(self, van_kan, pprime, p, bc_src=None):
        _check(lib().pp_update_pressure_device(self.h, int(van_kan), pprime.data_ptr(), p.data_ptr(),
                                               None if bc_src is None else bc_src.data_ptr()))

    # ---- implicit momentum solve ----
    def set_fresh_cells(self, tables):
        """pp_set_fresh_cells from the tables build_geometry returns as
        extras["fresh"] (keys i, j, k, ic, jc, kc, coeffD)."""
        a = [np.ascontiguousarray(tables[k], dtype=np.int32) for k in ("i", "j", "k", "ic", "jc", "kc")]
        cd = np.ascontiguousarray(tables["coeffD"], dtype=np.float64).reshape(-1, 8)
        n = len(a[0])
        assert all(len(x) == n for x in a) and len(cd) == n
        _check(lib().pp_set_fresh_cells(self.h, n, *[x.ctypes.data for x in a], cd.ctypes.data))

    def set_solve_ad(self, dt, visc_tot, bcxvisc, bcyvisc, bczvisc):
        """pp_set_solve_ad: the implicit viscous coefficients (set_solve_ad)."""
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (visc_tot, bcxvisc, bcyvisc, bczvisc)]
        for a in f:
            assert a.shape == self.shape
        _check(lib().pp_set_solve_ad(self.h, float(dt), *[a.ctypes.data for a in f]))

    def set_solve_ad_device(self, dt, visc_tot, bcxvisc, bcyvisc, bczvisc):
        _check(lib().pp_set_solve_ad_device(self.h, float(dt), *[a.data_ptr() for a in
                                                                 (visc_tot, bcxvisc, bcyvisc, bczvisc)]))

    def solve_ad(self, advec_scheme, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, vel_bi=None,
                 itermax_ad=1000, restol_ad=1e-10):
        """pp_solve_ad on copies of the host fields; returns (u, v, w, stats)."""
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (u, v, w)]
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (nlu, nlv, nlw, face_u, face_v, face_w)]
        for a in out + ins:
            assert a.shape == self.shape
        vb = None if vel_bi is None else np.ascontiguousarray(vel_bi, dtype=np.float64)
        st = ADStats()
        _check(lib().pp_solve_ad(self.h, int(advec_scheme), *[a.ctypes.data for a in out],
                                 *[a.ctypes.data for a in ins], None if vb is None else vb.ctypes.data,
                                 int(itermax_ad), float(restol_ad), C.byref(st)), allow=(PP_OK, PP_NOT_CONVERGED))
        return out[0], out[1], out[2], st.as_dict()

    def solve_ad_device(self, advec_scheme, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, vel_bi=None,
                        itermax_ad=1000, restol_ad=1e-10):
        """pp_solve_ad_device on torch float64 tensors (u, v, w updated in place); returns the stats."""
        st = ADStats()
        t = (u, v, w, nlu, nlv, nlw, face_u, face_v, face_w)
        _check(lib().pp_solve_ad_device(self.h, int(advec_scheme), *[a.data_ptr() for a in t],
                                        None if vel_bi is None else vel_bi.data_ptr(), int(itermax_ad),
                                        float(restol_ad), C.byref(st)), allow=(PP_OK, PP_NOT_CONVERGED))
        return st.as_dict()

    def solve_ad_cn2(self, u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, van_kan, p, bcxu, bcyv, bczw,
                     vel_bi=None, itermax_ad=1000, restol_ad=1e-10):
        """pp_solve_ad_cn2 on copies; returns (u, v, w, (face_u, face_v, face_w), stats)."""
        out = [np.array(a, dtype=np.float64, order="C", copy=True) for a in (u, v, w, face_u, face_v, face_w)]
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (nlu, nlv, nlw)]
        fv_in = [None if p is None else np.ascontiguousarray(p, dtype=np.float64)] + \
                [np.ascontiguousarray(a, dtype=np.float64) for a in (bcxu, bcyv, bczw)]
        for a in out + ins + [x for x in fv_in if x is not None]:
            assert a.shape == self.shape
        vb = None if vel_bi is None else np.ascontiguousarray(vel_bi, dtype=np.float64)
        st = ADStats()
        _check(lib().pp_solve_ad_cn2(self.h, *[a.ctypes.data for a in out[:3]], *[a.ctypes.data for a in ins],
                                     *[a.ctypes.data for a in out[3:]], int(van_kan),
                                     *[None if a is None else a.ctypes.data for a in fv_in],
                                     None if vb is None else vb.ctypes.data, int(itermax_ad), float(restol_ad),
                                     C.byref(st)), allow=(PP_OK, PP_NOT_CONVERGED))
        return out[0], out[1], out[2], tuple(out[3:]), st.as_dict()

    def face_vel(self, van_kan, dt, p, u, v, w, bcxu, bcyv, bczw, faces=None):
        """pp_face_vel: returns (face_u, face_v, face_w), written over copies of
        `faces` (zeros when None)."""
        ins = [None if p is None else np.ascontiguousarray(p, dtype=np.float64)] + \
              [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, bcxu, bcyv, bczw)]
        out = [np.zeros(self.shape) if faces is None else np.array(f, dtype=np.float64, order="C", copy=True)
               for f in (faces or (None, None, None))]
        _check(lib().pp_face_vel(self.h, int(van_kan), float(dt), *[None if a is None else a.ctypes.data for a in ins],
                                 *[a.ctypes.data for a in out]))
        return out

    def calc_residual_ad(self, advec_scheme, var, r, face_u, face_v, face_w):
        """pp_calc_residual_ad: returns (max |res|, 1-based loc, res field)."""
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (var, r, face_u, face_v, face_w)]
        for a in f:
            assert a.shape == self.shape
        res = np.zeros(self.shape)
        m = C.c_double()
        loc = (C.c_int32 * 3)()
        _check(lib().pp_calc_residual_ad(self.h, int(advec_scheme), *[a.ctypes.data for a in f], res.ctypes.data,
                                         C.byref(m), loc))
        return m.value, tuple(loc), res

    def divergence(self, face_u, face_v, face_w, dtinv=1.0):
        """rhs_poisson (non-cut-cell): returns (nlu0, sum(nlu0*dV))."""
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in (face_u, face_v, face_w)]
        for a in f:
            assert a.shape == self.shape
        out = np.zeros(self.shape)
        s = C.c_double()
        _check(lib().pp_divergence(self.h, f[0].ctypes.data, f[1].ctypes.data, f[2].ctypes.data, float(dtinv),
                                   out.ctypes.data, C.byref(s)))
        return out, s.value

    def divergence_noalloc(self, face_u, face_v, face_w, nlu0, dtinv=1.0, want_sum=True):
        """pp_divergence_noalloc on torch float64 device tensors (nlu0 written
        in place); returns sum(nlu0*dV), or None with want_sum=False (the call
        then returns without synchronising)."""
        s = C.c_double()
        _check(lib().pp_divergence_noalloc(self.h, face_u.data_ptr(), face_v.data_ptr(), face_w.data_ptr(),
                                           float(dtinv), nlu0.data_ptr(), C.byref(s) if want_sum else None))
        return s.value if want_sum else None

    def mean_remove(self, p):
        m = C.c_double()
        _check(lib().pp_mean_remove(self.h, p.ctypes.data, C.byref(m)))
        return m.value

    def null_vector(self):
        w = np.zeros(self.shape)
        _check(lib().pp_null_vector(self.h, w.ctypes.data))
        return w

    def level_info(self):
        n = C.c_int32()
        dims = (C.c_int32 * 48)()
        _check(lib().pp_level_info(self.h, C.byref(n), dims))
        return [(dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]) for l in range(n.value)]

    def setup_info(self):
        """Phase times (ms) of the last set_geometry and the null vector's
        iterations and relative residual."""
        ms = (C.c_double * 4)()
        it, res = C.c_int32(), C.c_double()
        _check(lib().pp_setup_info(self.h, ms, C.byref(it), C.byref(res)))
        return {"flags_ms": ms[0], "ghost_tables_ms": ms[1], "coarse_ms": ms[2], "null_vector_ms": ms[3],
                "null_iterations": it.value, "null_residual": res.value}

    def time_kernel(self, kernel_class, level=0, precision=PP_PREC_MIXED, reps=50):
        """(avg_us, bytes_per_launch) of reps back-to-back launches of one
        kernel class ("xline", "yline", "zline", "residual") on a level."""
        cls = KERNEL_CLASSES.index(kernel_class)
        us, b = C.c_double(), C.c_double()
        _check(lib().pp_time_kernel(self.h, cls, level, precision, reps, C.byref(us), C.byref(b)))
        return us.value, b.value

    def level_data(self, level):
        """pp_level_data: (flags (cz, cy, cx) uint8, {w, xc, cm, cp: [x, y, z] arrays})."""
        cx, cy, cz = self.level_info()[level]
        fl = np.zeros((cz, cy, cx), np.uint8)
        arrs = {k: np.zeros(cx + cy + cz) for k in ("w", "xc", "cm", "cp")}
        _check(lib().pp_level_data(self.h, level, fl.ctypes.data, *[arrs[k].ctypes.data for k in ("w", "xc", "cm", "cp")]))
        return fl, {k: [v[:cx], v[cx:cx + cy], v[cx + cy:]] for k, v in arrs.items()}

    def level_apply(self, op, level, a, b=None, precision=PP_PREC_FP64, omega=1.0):
        """pp_level_apply on dense level arrays (cz, cy, cx); returns out."""
        dims = self.level_info()
        lo = level + 1 if op in (PP_LOP_RESTRICT, PP_LOP_RESRESTRICT) else level
        cx, cy, cz = dims[lo]
        a = np.ascontiguousarray(a, dtype=np.float64)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros((cz, cy, cx))
        _check(lib().pp_level_apply(self.h, int(op), int(level), int(precision), float(omega), a.ctypes.data,
                                    None if bb is None else bb.ctypes.data, out.ctypes.data))
        return out

    def smooth(self, var, rhs, smoother=PP_SMOOTH_POINT_JACOBI, omega=1.0, sweeps=1):
        """pp_smooth: sweeps of the finest-level smoother on a copy of var
        (ghost rows not updated); returns the smoothed field."""
        v = np.array(var, dtype=np.float64, order="C", copy=True)
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        assert v.shape == self.shape and r.shape == self.shape
        _check(lib().pp_smooth(self.h, int(smoother), float(omega), int(sweeps), v.ctypes.data, r.ctypes.data))
        return v

    def tdma_batched(self, a, b, c, r, method):
        a, b, c, r = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r))
        nl, n = a.shape
        u = np.zeros_like(a)
        _check(lib().pp_tdma_batched(self.h, a.ctypes.data, b.ctypes.data, c.ctypes.data,
                                     r.ctypes.data, u.ctypes.data, n, nl, method))
        return u


@dataclass
class Ellipsoid:
    """Canonical ELLIPSOID body (MAIN.F90:1140-1168): centre, semi-axes
    (radiusx, radiusy, radiusz), rotation alpha about z in degrees and the
    marker count per latitude ring n_theta."""
    center: tuple
    semi: tuple
    alpha_deg: float = 0.0
    n_theta: int = 64


def _grid_desc(xn, yn, zn, device=0):
    g = GridDesc()
    g.nxc, g.nyc, g.nzc = len(xn) - 1, len(yn) - 1, len(zn) - 1
    g.x, g.y, g.z = _ptr(xn, C.c_double), _ptr(yn, C.c_double), _ptr(zn, C.c_double)
    for f in range(6):
        g.bc[f] = PP_BC_NEUMANN
    g.device = device
    return g


def build_geometry(xn, yn, zn, bodies, mode=PP_GEOM_ANALYTIC, probe_norm=2.0, iblank_prev=None, device=None):
    """set_boundary for canonical ellipsoids through pp_geom_build, or, with
    device = a GPU index (triangulated mode only), pp_geom_build_device.

    Returns (iblank, fresh, GhostSet, extras) with iblank/fresh in the
    reference layout and extras = dict(bi=..., ip=..., nrm=...)."""
    L = lib()
    xn, yn, zn = (np.ascontiguousarray(a, dtype=np.float64) for a in (xn, yn, zn))
    g = _grid_desc(xn, yn, zn, 0 if device is None else device)
    if device is not None and mode != PP_GEOM_TRIANGULATED:
        raise ValueError("the device build is the triangulated pipeline (mode=PP_GEOM_TRIANGULATED)")
    arr = (BodyDesc * max(len(bodies), 1))()
    for q, b in enumerate(bodies):
        arr[q].shape = PP_BODY_ELLIPSOID
        for d in range(3):
            arr[q].center[d] = b.center[d]
            arr[q].semi[d] = b.semi[d]
        arr[q].alpha_deg = b.alpha_deg
        arr[q].n_theta = b.n_theta
    prev = None if iblank_prev is None else np.ascontiguousarray(iblank_prev, dtype=np.int8)
    h = C.c_void_p()
    if device is None:
        _check(L.pp_geom_build(C.byref(h), C.byref(g), len(bodies), arr, mode, probe_norm,
                               None if prev is None else prev.ctypes.data))
    else:
        _check(L.pp_geom_build_device(C.byref(h), C.byref(g), len(bodies), arr, probe_norm,
                                      None if prev is None else prev.ctypes.data))
    return _geometry_result(h, xn, yn, zn)


def build_geometry_surfaces(xn, yn, zn, surfaces, probe_norm=2.0, iblank_prev=None, device=None):
    """set_boundary for UNSTRUCTURED_SURFACE bodies (pp_geom_build_surfaces,
    or pp_geom_build_surfaces_device on GPU `device`): surfaces = [(xyz
    (n_markers, 3), tri (n_elems, 3) 1-based, outside point)].  Returns what
    build_geometry returns."""
    L = lib()
    xn, yn, zn = (np.ascontiguousarray(a, dtype=np.float64) for a in (xn, yn, zn))
    g = _grid_desc(xn, yn, zn, 0 if device is None else device)
    keep = []
    arr = (SurfaceDesc * max(len(surfaces), 1))()
    for q, (xyz, tri, out) in enumerate(surfaces):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
        keep += [xyz, tri]
        arr[q].n_markers, arr[q].n_elems = len(xyz), len(tri)
        arr[q].xyz, arr[q].tri = xyz.ctypes.data, tri.ctypes.data
        for d in range(3):
            arr[q].outside[d] = out[d]
    prev = None if iblank_prev is None else np.ascontiguousarray(iblank_prev, dtype=np.int8)
    h = C.c_void_p()
    build = L.pp_geom_build_surfaces if device is None else L.pp_geom_build_surfaces_device
    _check(build(C.byref(h), C.byref(g), len(surfaces), arr, probe_norm, None if prev is None else prev.ctypes.data))
    return _geometry_result(h, xn, yn, zn)


def canonical_surface(body):
    """The marker surface PP_GEOM_TRIANGULATED builds for one Ellipsoid:
    (xyz (n_markers, 3), tri (n_elems, 3), 1-based marker ids)."""
    L = lib()
    d = BodyDesc()
    d.shape = PP_BODY_ELLIPSOID
    for q in range(3):
        d.center[q], d.semi[q] = body.center[q], body.semi[q]
    d.alpha_deg, d.n_theta = body.alpha_deg, body.n_theta
    nm, ne = C.c_int32(0), C.c_int32(0)
    _check(L.pp_geom_canonical_surface(C.byref(d), C.byref(nm), C.byref(ne), None, None))
    xyz, tri = np.zeros((nm.value, 3)), np.zeros((ne.value, 3), np.int32)
    _check(L.pp_geom_canonical_surface(C.byref(d), C.byref(nm), C.byref(ne), xyz.ctypes.data, tri.ctypes.data))
    return xyz, tri


def _geometry_result(h, xn, yn, zn):
    L = lib()
    try:
        ng, nb, nf = C.c_int32(), C.c_int32(), C.c_int32()
        _check(L.pp_geom_counts(h, C.byref(ng), C.byref(nb), C.byref(nf)))
        n = ng.value
        shape = (len(zn) + 2, len(yn) + 2, len(xn) + 2)
        ib = np.zeros(shape, np.int8)
        fr = np.zeros(shape, np.int8)
        ints = [np.zeros(max(n, 1), np.int32) for _ in range(6)]
        coeff = np.zeros((max(n, 1), 8))
        probe = np.zeros(max(n, 1))
        bi, ip, nrm = (np.zeros((max(n, 1), 3)) for _ in range(3))
        _check(L.pp_geom_get(h, ib.ctypes.data, fr.ctypes.data, *[a.ctypes.data for a in ints],
                             coeff.ctypes.data, probe.ctypes.data, bi.ctypes.data, ip.ctypes.data,
                             nrm.ctypes.data))
        body, elem = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        _check(L.pp_geom_ghost_elements(h, body.ctypes.data, elem.ctypes.data))
        coeffD = np.zeros((max(n, 1), 8))
        _check(L.pp_geom_dirichlet(h, coeffD.ctypes.data))
        fresh, fresh_error = _fresh_tables(h, nf.value)
    finally:
        L.pp_geom_destroy(h)
    gs = GhostSet(*[a[:n] for a in ints], coeff[:n], probe[:n])
    return ib, fr, gs, {"bi": bi[:n], "ip": ip[:n], "nrm": nrm[:n], "body": body[:n], "elem": elem[:n],
                        "coeffD": coeffD[:n], "fresh": fresh, "fresh_error": fresh_error}


FRESH_KEYS = ("i", "j", "k", "ic", "jc", "kc", "coeffD", "bi", "body", "elem")


def _fresh_tables(h, n):
    """pp_geom_fresh: (tables, None), or (None, message) when a fresh cell's
    stencil fails (the pressure path does not need the tables)."""
    m = max(n, 1)
    t = {k: np.zeros(m, np.int32) for k in ("i", "j", "k", "ic", "jc", "kc", "body", "elem")}
    t["coeffD"] = np.zeros((m, 8))
    t["bi"] = np.zeros((m, 3))
    rc = lib().pp_geom_fresh(h, *[t[k].ctypes.data for k in FRESH_KEYS])
    if rc != PP_OK:
        return None, lib().pp_last_error().decode()
    return {k: v[:n] for k, v in t.items()}, None


def tanh_nodes(n, length, beta):
    """Nodes clustered at the domain centre: x(xi) = L/2 [1 + artanh((2xi-1) tanh beta)/beta],
    xi = j/n (beta = 0: uniform)."""
    xi = np.arange(n + 1) / n
    if beta == 0:
        return xi * length
    return 0.5 * length * (1.0 + np.arctanh((2.0 * xi - 1.0) * np.tanh(beta)) / beta)
