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
  body (BOUNDARY_MARKER_INIT.F90:209-248), for pp.build_geometry_surfaces;
* the unformatted restart pair restart_flow_out{1,2}.dat and
  restart_body_out{1,2}.dat (write_restart, MAIN.F90:2035-2191; read back by
  read_restart_flow / read_restart_body, :1665-2031): the flow fields and the
  body state, iblank and fresh cells, record by record.

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
import struct

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
# canonical_body_in.dat (read_inputs_bound, MAIN.F90:915-1300)
# ---------------------------------------------------------------------------
ELLIPSOID = 3
UNSTRUCTURED_SURFACE = 4
GCM_METHOD = 2
_MOTION_RECORDS = [
    "vxcentTrans:r vycentTrans:r vzcentTrans:r", "ampx:r ampy:r ampz:r", "freqx:r freqy:r freqz:r",
    "angvx:r angvy:r angvz:r", "phase:r", "ampangx:r ampangy:r ampangz:r", "freqangx:r freqangy:r freqangz:r",
    "density_fluid:r density_solid:r", "xcentConstr:r ycentConstr:r zcentConstr:r",
]
# per canonical_body_type, the records after wall_type (MAIN.F90:1118-1212)
_BODY_RECORDS = {
    1: ["nPtsBodyMarker:i dummy:i", "radiusx:r radiusy:r", "xcent:r ycent:r", "alpha:r"] + _MOTION_RECORDS,
    2: ["nPtsBodyMarker:i dummy:i", "radiusx:r radiusy:r", "xcent:r ycent:r", "alpha:r"] + _MOTION_RECORDS,
    ELLIPSOID: ["n_theta:i dummy:i", "radiusx:r radiusy:r radiusz:r", "xcent:r ycent:r zcent:r", "alpha:r"]
               + _MOTION_RECORDS,
    4: ["dof1:i dof2:i dof3:i dof4:i dof5:i dof6:i", "nPtsBodyMarker:i totNumTriElem:i",
        "rigidRef1:i rigidRef2:i rigidRef3:i", "dummy1:r dummy2:r dummy3:r", "xcent:r ycent:r zcent:r",
        "quat0:r quat1:r quat2:r quat3:r"] + _MOTION_RECORDS
       + ["hingemarker:i", "thickoverlength:r", "Limiter_angle_amp:r", "A_x:r A_theta:r",
          "FreqN_torsion:r Freq_prsb:r"],
}
_ELLIPSOID_RECORDS = _BODY_RECORDS[ELLIPSOID]


def _body_records(b):
    recs = list(_BODY_RECORDS[b["canonical_body_type"]])
    if b["canonical_body_type"] == 4 and b.get("body_dim", 3) > 2:
        recs.append("DepthOverLength:r")
    return recs


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
    """(header, bodies): the header record, one dict per body (ELLIPTIC_CYLINDER,
    GENERAL_CYLINDER, ELLIPSOID or UNSTRUCTURED_SURFACE records) and, in the
    header, the channel-flow gates (a list of dicts) and Prsb_MomentRef that
    follow the bodies."""
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
        if b["canonical_body_type"] not in _BODY_RECORDS:
            raise ValueError(f"body {n + 1}: unknown canonical_body_type {b['canonical_body_type']}")
        for rec in _body_records(b):
            read_record(lines, _spec(rec), b)
        if b["canonical_body_type"] == ELLIPSOID:
            b["n_phi"] = ellipsoid_n_phi(b["n_theta"], b["radiusx"], b["radiusz"], boundary_formulation)
            b["nPtsBodyMarker"] = b["n_theta"] * b["n_phi"]
        if b["canonical_body_type"] == UNSTRUCTURED_SURFACE:
            b["alpha"] = 0.0
        bodies.append(b)
    gates = []
    if head["channel_flow"]:
        ng = read_record(lines, _spec("nGate:i"))["nGate"]
        for _ in range(ng):
            gates.append(read_record(lines, _spec("bcTypeGate:i bcGateU:r bcGateV:r bcGateW:r freqGate:r")))
    head["gates"] = gates
    try:
        head.update(read_record(lines, _spec("Prsb_MomentRef:l")))
    except StopIteration:
        head["Prsb_MomentRef"] = False
    return head, bodies


def write_canonical_bodies(path, bodies, channel_flow=False, n_section=0, gates=(), prsb_moment_ref=False):
    """canonical_body_in.dat for solid bodies: dicts with canonical_body_type
    (default ELLIPSOID) and the records of its type (missing values are
    written as zeros; an ELLIPSOID needs n_theta, radiusx/y/z,
    xcent/ycent/zcent, alpha; an UNSTRUCTURED_SURFACE nPtsBodyMarker and
    totNumTriElem matching its unstruc_surface_in.dat block)."""
    out = [f"{len(bodies)} {len(bodies)} 0 {n_section} {'T' if channel_flow else 'F'} F"]
    for b in bodies:
        b = dict(b)
        b.setdefault("canonical_body_type", ELLIPSOID)
        b.setdefault("body_dim", 3)
        out.append(f"{b['canonical_body_type']} {b['body_dim']} {b.get('boundary_motion_type', 0)}")
        out.append(str(b.get("wall_type", 0)))
        recs = _body_records(b)
        vals = dict.fromkeys(n for rec in recs for n, _ in _spec(rec))
        vals.update({k: 0 for k in vals})
        vals.update(b)
        for rec in recs:
            out.append(" ".join(_fmt(vals[n], k) for n, k in _spec(rec)))
    if channel_flow:
        out.append(str(len(gates)))
        for g in gates:
            out.append(" ".join(_fmt(g[n], k) for n, k in
                                _spec("bcTypeGate:i bcGateU:r bcGateV:r bcGateW:r freqGate:r")))
    out.append("T" if prsb_moment_ref else "F")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def case_bodies(case_dir, vals=None):
    """The case's ellipsoids as pp.Ellipsoid (centre, semi-axes, alpha in
    degrees, n_theta); a case with other body types raises ValueError (use
    case_geometry)."""
    from . import pp
    vals = vals or read_input_dat(os.path.join(case_dir, "input.dat"))
    if vals.get("internal_boundary_present", 1) == 0:
        return []
    _, bodies = read_canonical_bodies(os.path.join(case_dir, "canonical_body_in.dat"),
                                      vals.get("boundary_formulation", GCM_METHOD))
    if any(b["canonical_body_type"] != ELLIPSOID for b in bodies):
        raise ValueError("the case has bodies other than ELLIPSOID: use case_geometry")
    return [pp.Ellipsoid((b["xcent"], b["ycent"], b["zcent"]), (b["radiusx"], b["radiusy"], b["radiusz"]),
                         b["alpha"], b["n_theta"]) for b in bodies]


def case_geometry(case_dir, vals=None, nodes=None, iblank_prev=None, probe_norm=None):
    """set_boundary of a case whose bodies are ELLIPSOID or UNSTRUCTURED_SURFACE
    (in any mix): the ellipsoids' marker surfaces (as the reference builds
    them, oriented against the origin) and the surfaces of
    unstruc_surface_in.dat (one block per UNSTRUCTURED_SURFACE body, in body
    order) through the triangulated pipeline (pp.build_geometry_surfaces).
    Returns what pp.build_geometry returns.  ELLIPTIC_CYLINDER and
    GENERAL_CYLINDER bodies (2-D) raise NotImplementedError."""
    from . import pp
    vals = vals or read_input_dat(os.path.join(case_dir, "input.dat"))
    nodes = nodes or case_nodes(case_dir, vals)
    head, bodies = read_canonical_bodies(os.path.join(case_dir, "canonical_body_in.dat"),
                                         vals.get("boundary_formulation", GCM_METHOD))
    lines = None
    surfaces = []
    for n, b in enumerate(bodies):
        t = b["canonical_body_type"]
        if t == ELLIPSOID:
            xyz, tri = pp.canonical_surface(pp.Ellipsoid((b["xcent"], b["ycent"], b["zcent"]),
                                                         (b["radiusx"], b["radiusy"], b["radiusz"]),
                                                         b["alpha"], b["n_theta"]))
            surfaces.append((xyz, tri, (0.0, 0.0, 0.0)))
        elif t == UNSTRUCTURED_SURFACE:
            if lines is None:
                with open(os.path.join(case_dir, "unstruc_surface_in.dat")) as f:
                    lines = iter(f.read().splitlines())
            xyz, tri, out = read_unstruc_surface(lines, head["zoneSeparate"], head["channel_flow"])
            if len(xyz) != b["nPtsBodyMarker"] or len(tri) != b["totNumTriElem"]:
                raise ValueError(f"body {n + 1}: canonical_body_in.dat and unstruc_surface_in.dat disagree "
                                 f"({b['nPtsBodyMarker']}/{b['totNumTriElem']} against {len(xyz)}/{len(tri)})")
            surfaces.append((xyz, tri, out))
        else:
            raise NotImplementedError(f"body {n + 1}: canonical_body_type {t} (2-D cylinder) is not built")
    probe = vals.get("probeLengthNormalized", 2.0) if probe_norm is None else probe_norm
    return pp.build_geometry_surfaces(*nodes, surfaces, probe_norm=probe, iblank_prev=iblank_prev)


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


# ---------------------------------------------------------------------------
# unformatted sequential restart files (write_restart, MAIN.F90:2035-2191)
# ---------------------------------------------------------------------------
# A record is written as a 4-byte little-endian length, the data and the
# length again.  A record longer than SUBRECORD_MAX bytes is split into
# subrecords; the leading marker of a subrecord that another one follows and
# the trailing marker of a subrecord that continues another one are negated
# (the layout gfortran documents for its unformatted sequential files and
# writes compatibly with ifort).  Files of the reference's builds with other
# record-marker options (8-byte markers, big-endian) are not read.
SUBRECORD_MAX = 2 ** 31 - 9

# the flow record, in write order: 3-D fields (0:nx+1,0:ny+1,0:nz+1) and the
# 2-D outer pressure gradients pgradx1/2 (0:ny+1,0:nz+1), pgrady1/2
# (0:nx+1,0:nz+1), pgradz1/2 (0:nx+1,0:ny+1); REAL(8) throughout
FLOW_FIELDS = ["u", "v", "w", "face_u", "face_v", "face_w", "p", "nluold", "nlvold", "nlwold",
               "bcxu", "bcxv", "bcxw", "bcyu", "bcyv", "bcyw", "bczu", "bczv", "bczw",
               "pgradx1", "pgradx2", "pgrady1", "pgrady2", "pgradz1", "pgradz2",
               "viscTot", "bcxVisc", "bcyVisc", "bczVisc"]
ELLIPTIC_CYLINDER, GENERAL_CYLINDER = 1, 2


def flow_field_shape(name, nx, ny, nz):
    """numpy shape of a flow-record field for node counts nx, ny, nz."""
    if name.startswith("pgradx"):
        return (nz + 2, ny + 2)
    if name.startswith("pgrady"):
        return (nz + 2, nx + 2)
    if name.startswith("pgradz"):
        return (ny + 2, nx + 2)
    return (nz + 2, ny + 2, nx + 2)


def write_unformatted_record(f, *items):
    """One unformatted sequential record holding the items' bytes in order
    (numpy arrays in Fortran order: pass C-ordered arrays whose last axis is
    the Fortran first index)."""
    data = b"".join(np.ascontiguousarray(a).tobytes() for a in items)
    n = len(data)
    if n <= SUBRECORD_MAX:
        f.write(struct.pack("<i", n) + data + struct.pack("<i", n))
        return
    pos = 0
    while pos < n:
        m = min(SUBRECORD_MAX, n - pos)
        lead = -m if pos + m < n else m
        trail = -m if pos > 0 else m
        f.write(struct.pack("<i", lead) + data[pos:pos + m] + struct.pack("<i", trail))
        pos += m


def read_unformatted_record(f):
    """The data bytes of the next record (subrecords joined); EOFError at the
    end of the file."""
    parts = []
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise EOFError("no record left")
        lead = struct.unpack("<i", head)[0]
        m = abs(lead)
        data = f.read(m)
        tail = f.read(4)
        if len(data) != m or len(tail) != 4 or abs(struct.unpack("<i", tail)[0]) != m:
            raise ValueError("malformed unformatted record")
        parts.append(data)
        if lead >= 0:
            return b"".join(parts)


class _Rec:
    """Sequential reader of the items of one record."""

    def __init__(self, data):
        self.data, self.pos = data, 0

    def take(self, dtype, count=1, shape=None):
        dt = np.dtype(dtype)
        n = dt.itemsize * count
        if self.pos + n > len(self.data):
            raise ValueError("record shorter than its items")
        a = np.frombuffer(self.data, dtype=dt, count=count, offset=self.pos).copy()
        self.pos += n
        return a.reshape(shape) if shape is not None else a

    def done(self):
        if self.pos != len(self.data):
            raise ValueError(f"record has {len(self.data) - self.pos} bytes beyond its items")


def _rec(f):
    return _Rec(read_unformatted_record(f))


def write_restart_flow(path, nx, ny, nz, ntime, time, dt, re_, fields):
    """restart_flow_out*.dat: nx, ny, nz (node counts, INTEGER); ntime
    (INTEGER), time, dt, re (REAL(8)); the flow record (FLOW_FIELDS; a field
    left out of `fields` is written as zeros)."""
    arrays = []
    for name in FLOW_FIELDS:
        shp = flow_field_shape(name, nx, ny, nz)
        a = fields.get(name)
        a = np.zeros(shp) if a is None else np.asarray(a, dtype="<f8")
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape}, expected {shp}")
        arrays.append(a)
    with open(path, "wb") as f:
        write_unformatted_record(f, np.array([nx, ny, nz], "<i4"))
        write_unformatted_record(f, np.array([ntime], "<i4"), np.array([time, dt, re_], "<f8"))
        write_unformatted_record(f, *arrays)


def read_restart_flow(path):
    """read_restart_flow (MAIN.F90:1665-1754): a dict with nx, ny, nz, ntime,
    time, dt, re and `fields` (name -> array, FLOW_FIELDS order)."""
    with open(path, "rb") as f:
        r = _rec(f)
        nx, ny, nz = (int(v) for v in r.take("<i4", 3))
        r.done()
        r = _rec(f)
        ntime = int(r.take("<i4")[0])
        time_, dt, re_ = (float(v) for v in r.take("<f8", 3))
        r.done()
        r = _rec(f)
        fields = {}
        for name in FLOW_FIELDS:
            shp = flow_field_shape(name, nx, ny, nz)
            fields[name] = r.take("<f8", int(np.prod(shp)), shp)
        r.done()
    return {"nx": nx, "ny": ny, "nz": nz, "ntime": ntime, "time": time_, "dt": dt, "re": re_, "fields": fields}


# per body, in write order: (key, numpy dtype, count); the marker and
# element records follow
_BODY_HEAD = [
    [("npts_orig", "<i4", 1), ("npts", "<i4", 1)],
    [("canonical_body_type", "<i4", 1), ("body_dim", "<i4", 1), ("boundary_motion_type", "<i4", 1)],
    [("wall_type", "<i4", 1)],
    [("radius", "<f8", 3)], [("cent", "<f8", 3)], [("alpha", "<f8", 1)], [("vcent_trans", "<f8", 3)],
    [("amp", "<f8", 3)], [("freq", "<f8", 3)], [("angv", "<f8", 3)], [("phase", "<f8", 1)],
    [("ampang", "<f8", 3)], [("freqang", "<f8", 3)],
]


def _has_elements(t):
    return t in (ELLIPTIC_CYLINDER, GENERAL_CYLINDER, UNSTRUCTURED_SURFACE)


def write_restart_body(path, ntime, time, dt, re_, bodies, iblank, fresh, body_num=None, n_fresh=None):
    """restart_body_out*.dat: ntime, time, dt, re; nBody; per body the head
    records (_BODY_HEAD), the markers x(1:n), y(1:n), z(1:n), for cylinder and
    UNSTRUCTURED_SURFACE bodies the element count, triElemNeig(1,:),
    triElemNeig(2,:), triElemNeig(3,:) (1-based marker ids) and normDirFlag,
    and the marker velocities; then iblank and fresh_cell (INTEGER(1),
    reference layout), bodyNum (INTEGER, GCM only: pass body_num) and nFresh.
    A body is a dict with the _BODY_HEAD keys, "markers" (n, 3),
    "marker_vel" (n, 3) and, when it has elements, "tri" (ne, 3) and
    "norm_dir_flag"."""
    with open(path, "wb") as f:
        write_unformatted_record(f, np.array([ntime], "<i4"), np.array([time, dt, re_], "<f8"))
        write_unformatted_record(f, np.array([len(bodies)], "<i4"))
        for b in bodies:
            for rec in _BODY_HEAD:
                items = []
                for key, dt_, cnt in rec:
                    v = np.atleast_1d(np.asarray(b[key], dtype=dt_))
                    if v.size != cnt:
                        raise ValueError(f"body {key}: {v.size} values, expected {cnt}")
                    items.append(v)
                write_unformatted_record(f, *items)
            m = np.asarray(b["markers"], dtype="<f8").reshape(-1, 3)
            if len(m) != int(b["npts"]):
                raise ValueError("markers do not match npts")
            write_unformatted_record(f, m[:, 0], m[:, 1], m[:, 2])
            if _has_elements(int(b["canonical_body_type"])):
                t = np.asarray(b["tri"], dtype="<i4").reshape(-1, 3)
                write_unformatted_record(f, np.array([len(t)], "<i4"))
                write_unformatted_record(f, t[:, 0], t[:, 1], t[:, 2])
                write_unformatted_record(f, np.array([b["norm_dir_flag"]], "<f8"))
            mv = np.asarray(b["marker_vel"], dtype="<f8").reshape(-1, 3)
            write_unformatted_record(f, mv[:, 0], mv[:, 1], mv[:, 2])
        ib = np.asarray(iblank, dtype=np.int8)
        fr = np.asarray(fresh, dtype=np.int8)
        if ib.shape != fr.shape:
            raise ValueError("iblank and fresh_cell shapes differ")
        write_unformatted_record(f, ib, fr)
        if body_num is not None:
            write_unformatted_record(f, np.asarray(body_num, dtype="<i4"))
        nf = int((fr == 1).sum()) if n_fresh is None else int(n_fresh)
        write_unformatted_record(f, np.array([nf], "<i4"))


def read_restart_body(path, nx, ny, nz, gcm=True):
    """read_restart_body (MAIN.F90:1758-2031) for node counts nx, ny, nz: a
    dict with ntime, time, dt, re, bodies (as write_restart_body takes them),
    iblank, fresh, body_num (gcm only) and n_fresh."""
    shp = (nz + 2, ny + 2, nx + 2)
    n3 = int(np.prod(shp))
    with open(path, "rb") as f:
        r = _rec(f)
        ntime = int(r.take("<i4")[0])
        time_, dt, re_ = (float(v) for v in r.take("<f8", 3))
        r.done()
        r = _rec(f)
        nbody = int(r.take("<i4")[0])
        r.done()
        bodies = []
        for _ in range(nbody):
            b = {}
            for rec in _BODY_HEAD:
                r = _rec(f)
                for key, dt_, cnt in rec:
                    v = r.take(dt_, cnt)
                    b[key] = v if cnt > 1 else v[0].item()
                r.done()
            n = int(b["npts"])
            r = _rec(f)
            b["markers"] = r.take("<f8", 3 * n, (3, n)).T.copy()
            r.done()
            if _has_elements(int(b["canonical_body_type"])):
                r = _rec(f)
                ne = int(r.take("<i4")[0])
                r.done()
                r = _rec(f)
                b["tri"] = r.take("<i4", 3 * ne, (3, ne)).T.copy()
                r.done()
                r = _rec(f)
                b["norm_dir_flag"] = float(r.take("<f8")[0])
                r.done()
            r = _rec(f)
            b["marker_vel"] = r.take("<f8", 3 * n, (3, n)).T.copy()
            r.done()
            bodies.append(b)
        r = _rec(f)
        ib = r.take("i1", n3, shp)
        fr = r.take("i1", n3, shp)
        r.done()
        body_num = None
        if gcm:
            r = _rec(f)
            body_num = r.take("<i4", n3, shp)
            r.done()
        r = _rec(f)
        nfr = int(r.take("<i4")[0])
        r.done()
    return {"ntime": ntime, "time": time_, "dt": dt, "re": re_, "bodies": bodies, "iblank": ib, "fresh": fr,
            "body_num": body_num, "n_fresh": nfr}
