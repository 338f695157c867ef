"""AdH model files: .geo (mesh), .physics / .cover (materials and boundary strings), .params (models,
solver settings, series, boundary conditions), .init (initial state) and .design (grid, time control,
coupled models).

The .geo cards are read by libadhk.so (geo_read.c); the small control files are parsed here.  Card
layouts follow the reference's readers: sgrid_read.c:84-130 / :540-617 (ND, TRI, SEG),
smodel_super_read.c:48-64 and :143-159 (MODEL, material and string maps), read_bc_MODEL.c:29-153
(BC <string> <model> <type> <kind> <series>), the SERIES cards of sseries (time / value rows), the
IP cards (NTL, ITL, NIT, MIT) and the named datasets of the .init file (DEPTH, DA_VELOCITY, ...,
each closed by ENDDS).

load_sw2_model builds the Mesh, the initial state, the solver settings and the series bindings of
an SW2 model.  Only SW2 is on this build's path: a model whose materials name another physics (DW,
TRNS, ...) is refused with a message that says which.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .sw2 import Mesh


def read_geo(path):
    """(x, y, z, tri, tri_flag, seg, seg_flag) of a .geo file, 0-based, elements in file order."""
    lib = L.load()
    p = os.fsencode(path)
    n, t, s, o = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rc = lib.adhk_geo_counts(p, C.byref(n), C.byref(t), C.byref(s), C.byref(o))
    if rc != 0:
        raise L.AdhkError(f"{path}: cannot count the mesh cards (code {rc})")
    if o.value:
        raise L.AdhkError(f"{path}: {o.value} QUAD/TET/PRISM elements; the SW2 step takes triangles only")
    N, E, S = n.value, t.value, s.value
    x, y, z = np.zeros(N), np.zeros(N), np.zeros(N)
    tri, tfl = np.zeros((E, 3), np.int32), np.zeros(E, np.int32)
    seg, sfl = np.zeros((max(S, 1), 2), np.int32), np.zeros(max(S, 1), np.int32)
    rc = lib.adhk_geo_read(p, N, L.dptr(x), L.dptr(y), L.dptr(z), E, L.iptr(tri), L.iptr(tfl), S, L.iptr(seg),
                           L.iptr(sfl))
    if rc != 0:
        where = f" at line {-rc - 2}" if rc < -1 else ""
        raise L.AdhkError(f"{path}: malformed mesh card{where}")
    return x, y, z, tri, tfl, seg[:S], sfl[:S]


def _cards(path):
    with open(path) as f:
        for raw in f:
            line = raw.split("!")[0].split("#")[0].strip()
            if line:
                yield line.split()


def read_physics(path, ntri, nseg):
    """Material of each triangle and boundary string of each segment (TRI <mat> / SEG <string> lines, in
    element order; 0-based)."""
    tri_mat, seg_str = [], []
    for c in _cards(path):
        if c[0] == "TRI":
            tri_mat.append(int(c[1]) - 1)
        elif c[0] == "SEG":
            seg_str.append(int(c[1]) - 1)
    if len(tri_mat) != ntri or len(seg_str) != nseg:
        raise ValueError(f"{path}: {len(tri_mat)} TRI / {len(seg_str)} SEG lines for {ntri} / {nseg} elements")
    return np.array(tri_mat, np.int32), np.array(seg_str, np.int32)


@dataclass
class Params:
    models: dict = field(default_factory=dict)     # material id (1-based) -> model name
    ip: dict = field(default_factory=dict)         # NTL, ITL, NIT, MIT
    series: dict = field(default_factory=dict)     # series id (1-based) -> (t, v[:, nvalues])
    bcs: list = field(default_factory=list)        # (string id 1-based, model, type, kind, series ids)
    test: list = field(default_factory=list)


def _read_series_block(lines, i, npts):
    t, v = [], []
    for k in range(npts):
        row = [float(a) for a in lines[i + k]]
        t.append(row[0])
        v.append(row[1:])
    return np.array(t), np.array(v)


def read_params(path):
    """MODEL, IP, SERIES and BC cards of a .params file."""
    lines = list(_cards(path))
    p = Params()
    i = 0
    while i < len(lines):
        c = lines[i]
        if c[0] == "MODEL":
            p.models[int(c[1])] = c[2]
        elif c[0] == "IP" and len(c) >= 3:
            p.ip[c[1]] = float(c[2])
        elif c[0] == "TEST":
            p.test = c[1:]
        elif c[0] == "SERIES" and len(c) >= 4:
            sid, npts = int(c[2]), int(c[3])
            p.series[sid] = _read_series_block(lines, i + 1, npts)
            i += npts
        elif c[0] == "BC" and len(c) >= 5:
            p.bcs.append((int(c[1]), c[2], c[3], c[4], [int(a) for a in c[5:]]))
        i += 1
    return p


def read_design(path):
    """GRID, COVERAGE, TC T0 / TF, SERIES DT and MONO / COUPLE cards of a .design file."""
    lines = list(_cards(path))
    d = {"grid": None, "coverage": None, "t0": 0.0, "tf": None, "dt_series": None, "mono": [], "couple": []}
    i = 0
    while i < len(lines):
        c = lines[i]
        if c[0] == "GRID":
            d["grid"] = c[1]
        elif c[0] == "COVERAGE":
            d["coverage"] = c[1]
        elif c[0] == "TC" and c[1] == "T0":
            d["t0"] = float(c[2])
        elif c[0] == "TC" and c[1] == "TF":
            d["tf"] = float(c[2])
        elif c[0] == "SERIES" and len(c) >= 4:
            npts = int(c[3])
            if c[1] == "DT":
                d["dt_series"] = _read_series_block(lines, i + 1, npts)
            i += npts
        elif c[0] == "MONO":
            d["mono"].append((int(c[1]), c[2]))
        elif c[0] == "COUPLE":
            d["couple"].append(c[1:])
        i += 1
    return d


def read_init(path, nnodes):
    """Named datasets of an .init file: {name: array (nnodes,) or (nnodes, k)}."""
    out, name, rows = {}, None, []
    with open(path) as f:
        for raw in f:
            s = raw.strip()
            if not s:
                continue
            if s == "ENDDS":
                a = np.array(rows, float)
                out[name] = a[:, 0] if a.ndim == 2 and a.shape[1] == 1 else a
                if len(a) != nnodes:
                    raise ValueError(f"{path}: dataset {name} has {len(a)} rows for {nnodes} nodes")
                name, rows = None, []
            elif name is None:
                name = s.split()[0]
            else:
                rows.append([float(v) for v in s.split()])
    return out


@dataclass
class SW2Model:
    mesh: Mesh
    sol0: np.ndarray
    dt: float
    nsteps: int
    solver: dict
    series: dict          # series id -> (t, v)
    bindings: list        # (segment indices, series id, scale)


def load_sw2_model(design_path, model=1, defaults=None):
    """Mesh, initial state and settings of model `model` (its MONO number) of an AdH design.

    BC kinds: NB VEL (inflow velocity along the inward normal, from a series), NB DIS / NB FLUX (a
    discharge or unit flux, explicit flow), NB OTW / DB OVH (tailwater depth).  `defaults` fills what
    the files leave out (Manning n per material, a series for a BC that names none)."""
    defaults = defaults or {}
    base = os.path.dirname(os.path.abspath(design_path))
    d = read_design(design_path)
    names = dict(d["mono"])
    if model not in names:
        raise ValueError(f"{design_path}: no MONO {model}")
    stem = os.path.join(base, names[model])
    x, y, z, tri, _, seg, _ = read_geo(os.path.join(base, d["grid"]))
    prm = read_params(stem + ".params")
    tri_mat, seg_str = read_physics(stem + ".physics", len(tri), len(seg))
    used = sorted(set(int(m) + 1 for m in tri_mat))
    other = {m: prm.models.get(m, "?") for m in used if prm.models.get(m) != "SW2"}
    if other:
        # The reference registers no residual for DW or TRNS materials: select_resid_func returns NULL
        # for them (src/structs/sadh_def/sadh_def.c:261-263, :267-269), and load_pdes calls the pointer
        # (src/residual/assemble_residual.c:20), so its own assembly cannot take such an element either.
        raise NotImplementedError(f"{stem}.params: materials {other} are not SW2; this build runs the SW2 "
                                  f"Newton step only (the reference has no residual function for DW or TRNS)")
    # triangles counter-clockwise
    a = (x[tri[:, 1]] - x[tri[:, 0]]) * (y[tri[:, 2]] - y[tri[:, 0]]) - \
        (x[tri[:, 2]] - x[tri[:, 0]]) * (y[tri[:, 1]] - y[tri[:, 0]])
    tri = tri.copy()
    tri[a < 0] = tri[a < 0][:, [0, 2, 1]]
    mann = np.array([defaults.get("manning", {}).get(int(m) + 1, 0.03) for m in tri_mat])
    # boundary elements that carry a condition, oriented with the domain on the left
    from .channel import boundary_edges
    be = {tuple(sorted(e)): tuple(e) for e in boundary_edges(tri)}
    segs, kinds, vals, binds = [], [], [], []
    series = dict(prm.series)
    for string, mdl, btype, kind, sids in prm.bcs:
        if mdl != "SW2":
            continue
        members = np.flatnonzero(seg_str == string - 1)
        for s_i in members:
            key = tuple(sorted(seg[s_i]))
            if key not in be:
                raise ValueError(f"segment {s_i + 1} of string {string} is not on the boundary")
            segs.append(be[key])
        first = len(segs) - len(members)
        idx = np.arange(first, len(segs))
        sid = sids[0] if sids else defaults.get("series", {}).get(string)
        if (btype, kind) == ("NB", "VEL"):
            kinds += [L.BC_VEL] * len(members)
            vals += [[0.0, 0.0]] * len(members)
            binds.append((idx, sid, "normal_velocity"))
        elif (btype, kind) in (("NB", "DIS"), ("NB", "FLUX")):
            kinds += [L.BC_FLUX] * len(members)
            vals += [[0.0, 0.0]] * len(members)
            binds.append((idx, sid, "discharge" if kind == "DIS" else "unit_flux"))
        elif (btype, kind) in (("NB", "OTW"), ("DB", "OVH")):
            kinds += [L.BC_DEPTH] * len(members)
            vals += [[0.0, 0.0]] * len(members)
            binds.append((idx, sid, "depth"))
        else:
            raise NotImplementedError(f"BC {btype} {kind} on string {string}")
    seg_arr = np.array(segs, np.int32).reshape(-1, 2)
    mesh = Mesh(x, y, z, np.ascontiguousarray(tri), seg_arr, np.array(kinds, np.int32),
                np.array(vals, float).reshape(-1, 2), mann)
    init = read_init(stem + ".init", len(x)) if os.path.exists(stem + ".init") else {}
    sol0 = np.zeros(3 * len(x))
    sol0[0::3] = init.get("DEPTH", np.full(len(x), defaults.get("depth", 1.0)))
    if "DA_VELOCITY" in init:
        sol0[1::3], sol0[2::3] = init["DA_VELOCITY"][:, 0], init["DA_VELOCITY"][:, 1]
    t_dt, v_dt = d["dt_series"] if d["dt_series"] is not None else (np.array([0.0]), np.array([[1.0]]))
    dt = float(v_dt[0][0])
    tf = d["tf"] if d["tf"] is not None else dt
    nsteps = int(round((tf - d["t0"]) / dt))
    solver = {"tol_nonlin": prm.ip.get("NTL", 1e-8), "max_nonlin_it": int(prm.ip.get("NIT", 10)),
              "lin_max_it": int(prm.ip.get("MIT", 100))}
    # scales per binding: a VEL segment takes the inward normal times the series value
    bindings = []
    for idx, sid, what in binds:
        if sid is None:
            raise ValueError(f"a boundary string has no series and no default: {what}")
        if what == "normal_velocity":
            s = seg_arr[idx]
            dx, dy = x[s[:, 1]] - x[s[:, 0]], y[s[:, 1]] - y[s[:, 0]]
            ln = np.hypot(dx, dy)
            nx, ny = dy / ln, -dx / ln  # outward normal (selem_1d.c:53-54)
            bindings.append((idx, sid, 0, -nx))
            bindings.append((idx, sid, 1, -ny))
        elif what == "discharge":
            s = seg_arr[idx]
            w = np.hypot(x[s[:, 1]] - x[s[:, 0]], y[s[:, 1]] - y[s[:, 0]]).sum()
            bindings.append((idx, sid, 0, np.full(len(idx), -1.0 / w)))
        elif what == "unit_flux":
            bindings.append((idx, sid, 0, np.full(len(idx), -1.0)))
        else:
            bindings.append((idx, sid, 0, np.ones(len(idx))))
    return SW2Model(mesh, sol0, dt, nsteps, solver, series, bindings)


def apply_bindings(sm, model: SW2Model):
    """Load the model's series into sm and bind its boundary segments (one binding per value)."""
    for sid, (t, v) in model.series.items():
        sm.set_series(sid, t, v[:, 0])
    for idx, sid, comp, scale in model.bindings:
        for i, s in zip(idx, scale):
            sm.bind_series([i], sid, float(s), component=comp)
