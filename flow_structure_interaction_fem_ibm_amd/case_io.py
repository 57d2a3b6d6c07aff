"""Readers and writers of the reference's case files, so a case set up for
the Fortran solver can drive this library (and a case made here can be
handed back to it):

* input.dat: the 120 list-directed records read by read_inputs
  (MAIN.F90:398-517), in order; a record with no items is a comment line;
* xgrid.dat / ygrid.dat / zgrid.dat: one "index coordinate" pair per node
  (CORE_GRID.F90:46-80), and the uniform-grid rule x(i) = (i-1) xout/(nx-1);
* canonical_body_in.dat: the header and the ELLIPSOID bodies of
  read_inputs_bound (MAIN.F90:915-1166), with the reference's marker count
  n_phi from n_theta and the aspect ratio;
* unstruc_surface_in.dat: one triangulated surface per UNSTRUCTURED_SURFACE
  body (BOUNDARY_MARKER_INIT.F90:209-248), for pp.build_geometry_surfaces.

Fortran list-directed input is restated for what these files use: values
separated by blanks and/or one comma, null values (two commas with only
blanks between, a comma opening a record, or r* with no value: the item
keeps its value), r*value repeat counts, a slash ending the record (later
items keep their value), D exponents, logical T/F/.TRUE./.FALSE.; a READ
takes further records until its items are filled and drops the rest of its
last record.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

# ---------------------------------------------------------------------------
# list-directed input
# ---------------------------------------------------------------------------
def _fields(line, after_value=False):
    """The values of one record in order: strings, None for a null value,
    "/" for a slash (the end of the READ); a "!" starts a comment.

    after_value carries the separator state of the READ across its records:
    the end of a record acts as a blank, so a record that opens with a comma
    after a record that ended in a value has one separator, not a null.
    Returns (values, after_value at the end of the record)."""
    out, i, n = [], 0, len(line)
    while i < n:
        ch = line[i]
        if ch in " \t":
            i += 1
        elif ch == ",":
            if not after_value:
                out.append(None)
            after_value = False
            i += 1
        elif ch == "/":
            out.append("/")
            break
        elif ch == "!":
            break
        else:
            j = i
            while j < n and line[j] not in " \t,/":
                j += 1
            tok = line[i:j]
            m = re.fullmatch(r"(\d+)\*(.*)", tok)
            if m:
                out.extend([m.group(2) or None] * int(m.group(1)))
            else:
                out.append(tok)
            after_value = True
            i = j
    return out, after_value


def _convert(tok, kind, name):
    t = tok.strip()
    if kind == "l":
        u = t.upper().lstrip(".")
        if u.startswith("T"):
            return True
        if u.startswith("F"):
            return False
        raise ValueError(f"{name}: {tok!r} is not a logical")
    if kind == "i":
        return int(t)
    if kind == "r":
        return float(t.replace("d", "e").replace("D", "E"))
    raise ValueError(kind)


def read_record(lines, items, values=None):
    """One list-directed READ of `items` ((name, kind) pairs, kind 'i', 'r'
    or 'l') from the line iterator; returns the dict of values read (items
    after a slash are left out, or keep their entry in `values`)."""
    out = {} if values is None else values
    want = list(items)
    if not want:
        next(lines)  # READ(u,*) with no list skips one record
        return out
    pos = 0
    after_value = False  # no value yet: a leading comma is a null
    while pos < len(want):
        toks, after_value = _fields(next(lines), after_value)
        for tok in toks:
            if tok == "/":
                return out
            name, kind = want[pos]
            if tok is not None:
                out[name] = _convert(tok, kind, name)
            pos += 1
            if pos >= len(want):
                break
    return out


def _fmt(v, kind):
    if kind == "l":
        return "T" if v else "F"
    if kind == "i":
        return str(int(v))
    return repr(float(v))


# ---------------------------------------------------------------------------
# input.dat (MAIN.F90:398-517): "" = a comment record
# ---------------------------------------------------------------------------
def _spec(s):
    return [tuple(x.split(":")) for x in s.split()] if s else []


_BCV = "{b}:i u{d}:r v{d}:r w{d}:r freq_u{d}:r freq_v{d}:r freq_w{d}:r"
INPUT_DAT = [_spec(s) for s in [
    "", "nread:i ntime_skip_check:i dryRun:l ibkOut:l",
    "", "ndim:i flow_type:i",
    "", "nx:i ny:i nz:i",
    "", "xgrid_unif:i xout:r", "", "ygrid_unif:i yout:r", "", "zgrid_unif:i zout:r",
    "", "uinit:r vinit:r winit:r vper:r",
    "", "", "", _BCV.format(b="bcx1", d="x1"),
    "", "", _BCV.format(b="bcx2", d="x2"),
    "", "", _BCV.format(b="bcy1", d="y1"),
    "", "", _BCV.format(b="bcy2", d="y2"),
    "", "", _BCV.format(b="bcz1", d="z1"),
    "", "", _BCV.format(b="bcz2", d="z2"),
    "", "", "", "pbcx1:i pppx1:r", "", "", "pbcx2:i pppx2:r", "", "", "pbcy1:i pppy1:r",
    "", "", "pbcy2:i pppy2:r", "", "", "pbcz1:i pppz1:r", "", "", "pbcz2:i pppz2:r",
    "", "", "no_tsteps:i nmonitor:i ndump:i nrestart:i nstat:i nmonitor_probe_liftdrag:i STATS_SUM:i",
    "", "format_dump:i",
    "", "re:r dt:r",
    "", "frac_step_type:i advec_scheme:i",
    "", "", "alfa:r",
    "", "", "internal_boundary_present:i",
    "", "body_type:i",
    "", "boundary_formulation:i",
    "", "probeLengthNormalized:r",
    "", "", "iterMax_ad:i restol_ad:r omega_adv:r ad_solver_type:i",
    "", "", "pp_solver_type:i iRedBlack:i",
    "", "omega:r",
    "", "iterMax_Poisson:i restol_Poisson:r iterResPoisson:i",
    "", "", "mgcyclex:i mgcycley:i mgcyclez:i",
    "", "iterFinest:i iterInter:i iterCoarsest:i",
    "", "mgLevels_X:i mgLevels_Y:i mgLevels_Z:i",
    "", "infoconv:i",
    "", "", "turbActive:i", "", "turbModel:i", "", "cSmagFix:r", "", "filterWidthModel:i",
    "", "testFilterDir_x:i testFilterDir_y:i testFilterDir_z:i",
    "", "homogDir_x:i homogDir_y:i homogDir_z:i",
    "", "fWidthRatio_x:r fWidthRatio_y:r fWidthRatio_z:r",
    "", "turbLagrTimeScale:r",
    "", "AdditionalOutput:i BinaryOutput:i",
    "", "Full_Coarsening:l cure_pressure_oscillations:l",
    "", "pressure_osc_velocity:l pressure_osc_pressure:l",
    "", "", "optimization:l",
    "", "evaluationInterval:i errorPermitted:r",
]]
assert len(INPUT_DAT) == 120
INPUT_KINDS = {n: k for rec in INPUT_DAT for n, k in rec}


def read_input_dat(path):
    """The values read_inputs takes from input.dat (MAIN.F90:398-517),
    keyed by the reference's variable names (array elements as _x/_y/_z)."""
    with open(path) as f:
        lines = iter(f.read().splitlines())
    vals = {}
    for q, rec in enumerate(INPUT_DAT):
        try:
            read_record(lines, rec, vals)
        except StopIteration:
            raise ValueError(f"{path}: ends before record {q + 1} of input.dat") from None
    return vals


def write_input_dat(path, values):
    """input.dat with every record in read_inputs' order; each comment
    record names the items of the next data record."""
    out = []
    for q, rec in enumerate(INPUT_DAT):
        if not rec:
            nxt = next((r for r in INPUT_DAT[q + 1:] if r), [])
            out.append("# " + " ".join(n for n, _ in nxt))
        else:
            out.append(" ".join(_fmt(values[n], k) for n, k in rec))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


# ---------------------------------------------------------------------------
# grid files (CORE_GRID.F90:36-80)
# ---------------------------------------------------------------------------
UNIFORM_GRID, NONUNIFORM_GRID = 1, 2


def read_grid_dat(path, n):
    """n node coordinates x(1..n) from "junk x(i)" records."""
    with open(path) as f:
        lines = iter(f.read().splitlines())
    x = np.empty(n)
    for i in range(n):
        rec = read_record(lines, [("junk", "r"), ("x", "r")])
        x[i] = rec["x"]
    return x


def write_grid_dat(path, x):
    with open(path, "w") as f:
        for i, v in enumerate(np.asarray(x, dtype=np.float64), start=1):
            f.write(f"{i} {float(v)!r}\n")


def case_nodes(case_dir, vals=None):
    """Node coordinates x(1..nx), y(1..ny), z(1..nz) of a case directory as
    make_grid forms them: uniform x(i) = (i-1) xout/(nx-1), else xgrid.dat."""
    vals = vals or read_input_dat(os.path.join(case_dir, "input.dat"))
    out = []
    for d, (n, unif, length, fname) in enumerate(
            ((vals["nx"], vals["xgrid_unif"], vals["xout"], "xgrid.dat"),
             (vals["ny"], vals["ygrid_unif"], vals["yout"], "ygrid.dat"),
             (vals["nz"], vals["zgrid_unif"], vals["zout"], "zgrid.dat"))):
        if unif == UNIFORM_GRID:
            delta = length / float(n - 1)
            out.append(np.array([float(i - 1) * delta for i in range(1, n + 1)]))
        else:
            out.append(read_grid_dat(os.path.join(case_dir, fname), n))
    return tuple(out)


# ---------------------------------------------------------------------------
# canonical_body_in.dat (MAIN.F90:915-1166), ELLIPSOID bodies
# ---------------------------------------------------------------------------
ELLIPSOID = 3
GCM_METHOD = 2
_ELLIPSOID_RECORDS = [
    "n_theta:i dummy:i", "radiusx:r radiusy:r radiusz:r", "xcent:r ycent:r zcent:r", "alpha:r",
    "vxcentTrans:r vycentTrans:r vzcentTrans:r", "ampx:r ampy:r ampz:r", "freqx:r freqy:r freqz:r",
    "angvx:r angvy:r angvz:r", "phase:r", "ampangx:r ampangy:r ampangz:r", "freqangx:r freqangy:r freqangz:r",
    "density_fluid:r density_solid:r", "xcentConstr:r ycentConstr:r zcentConstr:r",
]


def ellipsoid_n_phi(n_theta, radiusx, radiusz, boundary_formulation=GCM_METHOD):
    """n_phi of an ellipsoid's marker surface (MAIN.F90:1154-1165): the
    aspect ratio radiusz/radiusx with GCM, 1 otherwise; IDNINT rounds half
    away from zero."""
    aratio = radiusz / radiusx if boundary_formulation == GCM_METHOD else 1.0

    def idnint(v):
        return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))
    if n_theta % 2 == 0:
        return idnint(0.5 * n_theta * aratio)
    return idnint(0.5 * (n_theta - 1) * aratio) + 1


def read_canonical_bodies(path, boundary_formulation=GCM_METHOD):
    """(header, bodies): the header record and one dict per body; bodies
    other than ELLIPSOID raise NotImplementedError."""
    with open(path) as f:
        lines = iter(f.read().splitlines())
    head = read_record(lines, _spec("nbody:i nbody_solid:i nbody_membrane:i nSection:i channel_flow:l "
                                    "zoneSeparate:l"))
    bodies = []
    for n in range(head["nbody"]):
        spec = "canonical_body_type:i body_dim:i boundary_motion_type:i"
        if head["zoneSeparate"]:
            spec += " zoneMax:i"
        b = read_record(lines, _spec(spec))
        read_record(lines, _spec("wall_type:i"), b)
        if b["canonical_body_type"] != ELLIPSOID:
            raise NotImplementedError(f"body {n + 1}: canonical_body_type {b['canonical_body_type']} "
                                      f"(only ELLIPSOID = {ELLIPSOID} is read)")
        for rec in _ELLIPSOID_RECORDS:
            read_record(lines, _spec(rec), b)
        b["n_phi"] = ellipsoid_n_phi(b["n_theta"], b["radiusx"], b["radiusz"], boundary_formulation)
        bodies.append(b)
    return head, bodies


def write_canonical_bodies(path, bodies, channel_flow=False, n_section=0):
    """canonical_body_in.dat for resting solid ELLIPSOID bodies (dicts with
    n_theta, radiusx/y/z, xcent/ycent/zcent, alpha; the motion, density and
    constraint records are written as zeros)."""
    out = [f"{len(bodies)} {len(bodies)} 0 {n_section} {'T' if channel_flow else 'F'} F"]
    for b in bodies:
        out.append(f"{ELLIPSOID} 3 0")
        out.append("0")
        vals = dict.fromkeys(n for rec in _ELLIPSOID_RECORDS for n, _ in _spec(rec))
        vals.update({k: 0 for k in vals})
        vals.update(b)
        for rec in _ELLIPSOID_RECORDS:
            out.append(" ".join(_fmt(vals[n], k) for n, k in _spec(rec)))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def case_bodies(case_dir, vals=None):
    """The case's ellipsoids as pp.Ellipsoid (centre, semi-axes, alpha in
    degrees, n_theta)."""
    from . import pp
    vals = vals or read_input_dat(os.path.join(case_dir, "input.dat"))
    if vals.get("internal_boundary_present", 1) == 0:
        return []
    _, bodies = read_canonical_bodies(os.path.join(case_dir, "canonical_body_in.dat"),
                                      vals.get("boundary_formulation", GCM_METHOD))
    return [pp.Ellipsoid((b["xcent"], b["ycent"], b["zcent"]), (b["radiusx"], b["radiusy"], b["radiusz"]),
                         b["alpha"], b["n_theta"]) for b in bodies]


# ---------------------------------------------------------------------------
# unstruc_surface_in.dat (BOUNDARY_MARKER_INIT.F90:209-248), one body
# ---------------------------------------------------------------------------
def read_unstruc_surface(path_or_lines, zone_separate=False, channel_flow=False):
    """(xyz (n_markers, 3), tri (n_elems, 3) 1-based, outside point): a
    comment record, "nPtsBodyMarkerIn totNumTriElemIn", a comment record,
    "i x y z" per marker (plus zoneMarker or gateLabel when zoneSeparate or
    channel_flow), a comment record, "i n1 n2 n3" per element, a comment
    record, pointOutsideBody.  A line iterator continues with the next body."""
    lines = iter(open(path_or_lines).read().splitlines()) if isinstance(path_or_lines, str) else path_or_lines
    read_record(lines, [])
    h = read_record(lines, _spec("n:i ne:i"))
    read_record(lines, [])
    extra = " label:i" if (zone_separate or channel_flow) else ""
    xyz = np.empty((h["n"], 3))
    for m in range(h["n"]):
        r = read_record(lines, _spec("i:i x:r y:r z:r" + extra))
        xyz[m] = (r["x"], r["y"], r["z"])
    read_record(lines, [])
    tri = np.empty((h["ne"], 3), np.int32)
    for e in range(h["ne"]):
        r = read_record(lines, _spec("i:i a:i b:i c:i"))
        tri[e] = (r["a"], r["b"], r["c"])
    read_record(lines, [])
    o = read_record(lines, _spec("x:r y:r z:r"))
    return xyz, tri, (o["x"], o["y"], o["z"])


def write_unstruc_surface(path, xyz, tri, outside):
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    tri = np.asarray(tri).reshape(-1, 3)
    out = ["# nPtsBodyMarker totNumTriElem", f"{len(xyz)} {len(tri)}", "# markers"]
    out += [f"{m + 1} {x!r} {y!r} {z!r}" for m, (x, y, z) in enumerate(xyz.tolist())]
    out.append("# elements")
    out += [f"{e + 1} {a} {b} {c}" for e, (a, b, c) in enumerate(tri.tolist())]
    out.append("# pointOutsideBody")
    out.append(" ".join(repr(float(v)) for v in outside))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
