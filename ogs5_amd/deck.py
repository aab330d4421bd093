"""OGS-5 input decks: the subset that drives LIQUID_FLOW, GROUNDWATER_FLOW and
HEAT_TRANSPORT,
read into the scenarios the device path runs, and a driver for a deck's
time loop.

Files and keywords read (each as the reference's reader takes it):

  .pcs  #PROCESS $PCS_TYPE                       (FEM/rf_pcs.cpp PCSRead)
  .num  #NUMERICS $PCS_TYPE $LINEAR_SOLVER        (FEM/rf_num_new.cpp:526-552)
        $ELE_GAUSS_POINTS (:673) $NON_LINEAR_ITERATIONS (:412, relaxation)
  .tim  #TIME_STEPPING $PCS_TYPE $TIME_START $TIME_END $TIME_UNIT
        $TIME_STEPS "count dt" lines             (FEM/rf_tim_new.cpp:225-310)
  .msh  #FEM_MSH $NODES "id x y z", $ELEMENTS "id group [-1] type nodes"
                                                 (MSH/msh_mesh.cpp:437-491,
                                                  MSH/msh_elem.cpp:578)
  .gli  #POINTS "id x y z $NAME name", #POLYLINE $NAME $POINTS $EPSILON
  .bc   #BOUNDARY_CONDITION, .st #SOURCE_TERM, .ic #INITIAL_CONDITION:
        $PCS_TYPE $PRIMARY_VARIABLE $GEO_TYPE (POINT / POLYLINE / DOMAIN)
        $DIS_TYPE CONSTANT v                     (FEM/rf_bc_new.cpp:271-360)
  .mmp  #MEDIUM_PROPERTIES per material group: $GEOMETRY_DIMENSION
        $GEOMETRY_AREA $POROSITY 1 v $STORAGE 1 v
        $PERMEABILITY_TENSOR ISOTROPIC k | ORTHOTROPIC k.. | ANISOTROPIC k..
                                                 (FEM/rf_mmp_new.cpp:363-900)
  .msp  #SOLID_PROPERTIES $DENSITY 1 v, CAPACITY 1 v, CONDUCTIVITY 1 v
                                                 (FEM/rf_msp_new.cpp:171-300)
  .mfp  #FLUID_PROPERTIES $DENSITY 1 v $VISCOSITY 1 v
        $SPECIFIC_HEAT_CAPACITY 1 v $HEAT_CONDUCTIVITY 1 v $GRAVITY g
        $NON_GRAVITY                             (FEM/rf_mfp_new.cpp:312-840)
  .out  #OUTPUT $PCS_TYPE $NOD_VALUES $GEO_TYPE $DAT_TYPE $TIM_TYPE

Lines are read as GetLineFromFile1 does (FEM/files0.cpp:529): leading
blanks skipped, a line starting with ';' is a comment, text after ';' is
dropped; #STOP ends a file.

Coefficients (per element, from its material group):
  GROUNDWATER_FLOW  HEAD form: mass = storage, K = the conductivity tensor
                times tuf, no viscosity, no gravity (fem_ele_std.cpp:1625,
                2308-2384).
  LIQUID_FLOW   mass = storage (CalCoefMass, FEM/fem_ele_std.cpp:1580-1622,
                drho/dp = 0); K from the permeability tensor; mu = viscosity;
                gravity rho*g on the vertical axis when the mesh's coordinate
                flag ends in 2 (xz, xyz; MSH/msh_mesh.cpp:893-922) and the
                fluid has gravity (Assemble_Gravity :5859).
  HEAT_TRANSPORT  mass = n rho_f c_f + (1-n) |rho_s| c_s when a flow process
                runs, else |rho_s| c_s (HeatCapacity, rf_mmp_new.cpp:2720);
                conductivity (1-n) lambda_s + n lambda_f on the diagonal when
                a flow process runs (HeatConductivityTensor :2834); advection
                c_f rho_f with the flow's Gauss-point velocity.
  Time unit     MINUTE/HOUR/DAY/MONTH/YEAR -> time_unit_factor
                (FEM/rf_pcs.cpp:863-872).

Out of this subset (rejected with an error, not guessed): other process
types, curves and function-valued models, SURFACE geometries, distributed
source terms, other element types, several meshes.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib as L
from .mesh import EDIM, HEX8, LINE2, QUAD4, TET4, Mesh
from .workloads import Scenario, Step

GRAVITY_CONSTANT = 9.81  # FEM/rf_mfp_new.cpp:49
_ETYPES = {"line": LINE2, "quad": QUAD4, "hex": HEX8, "tet": TET4}
_PROCESS = {"LIQUID_FLOW": L.LIQUID_FLOW, "GROUNDWATER_FLOW": L.GROUNDWATER_FLOW,
            "HEAT_TRANSPORT": L.HEAT_TRANSPORT}
_FLOW = ("LIQUID_FLOW", "GROUNDWATER_FLOW")
_PRIMARY = {L.LIQUID_FLOW: "PRESSURE1", L.GROUNDWATER_FLOW: "HEAD",
            L.HEAT_TRANSPORT: "TEMPERATURE1"}
_TIME_UNIT = (("MINUTE", 60.0), ("HOUR", 3600.0), ("DAY", 86400.0),
              ("MONTH", 2592000.0), ("YEAR", 31536000.0))


class DeckError(ValueError):
    pass


# ---------------------------------------------------------------------------
# keyword blocks
# ---------------------------------------------------------------------------
def _data_lines(path: str) -> List[str]:
    out = []
    with open(path) as f:
        for raw in f:
            s = raw.strip()
            if not s or s.startswith(";"):
                continue
            j = s.find(";")
            if j >= 0:
                s = s[:j].rstrip()
            if s.startswith("#STOP"):
                break
            out.append(s)
    return out


def _blocks(path: str) -> List[Tuple[str, List[str]]]:
    """[(keyword, lines)] of a keyword file; lines keep their $SUBKEYWORDS."""
    blocks: List[Tuple[str, List[str]]] = []
    lines = _data_lines(path)
    cur = None
    for s in lines:
        if s.startswith("#"):
            cur = (s[1:].split()[0], [])
            blocks.append(cur)
        elif cur is not None:
            cur[1].append(s)
    return blocks


def _subkeys(lines: List[str]) -> List[Tuple[str, List[str]]]:
    """[(SUBKEY, data lines)] in file order."""
    out: List[Tuple[str, List[str]]] = []
    for s in lines:
        if s.startswith("$"):
            out.append((s[1:].split()[0], []))
        elif out:
            out[-1][1].append(s)
    return out


def _model_value(data: List[str], what: str) -> float:
    t = data[0].split() if data else []
    if len(t) < 2 or t[0] != "1":
        raise DeckError(f"{what}: only the constant model (1 value) is supported, got {data[:1]}")
    return float(t[1])


# ---------------------------------------------------------------------------
# mesh and geometry
# ---------------------------------------------------------------------------
def _msh_bulk(raw: bytes):
    """The $NODES and $ELEMENTS blocks parsed by numpy's C text parser when
    the mesh is plain (one mesh, one element type, no comments or $AREA
    lines inside the blocks); None otherwise (the line reader takes over)."""
    import warnings
    i0 = raw.find(b"#FEM_MSH")
    if i0 < 0 or raw.find(b"#FEM_MSH", i0 + 1) >= 0:
        return None

    def block(key):
        k = raw.find(key, i0)
        if k < 0:
            return None, None
        a = raw.index(b"\n", k) + 1
        b = raw.index(b"\n", a)
        count = int(raw[a:b].split()[0])
        ends = [e for e in (raw.find(b"$", b), raw.find(b"#", b)) if e >= 0]
        return count, raw[b:min(ends) if ends else len(raw)]

    n, nb = block(b"$NODES")
    E, eb = block(b"$ELEMENTS")
    if nb is None or eb is None or b";" in nb or b";" in eb:
        return None
    first = eb.split(b"\n", 2)[1].split() if E else []
    if len(first) < 4:
        return None
    j = 3 if first[2] == b"-1" else 2
    name = first[j].decode()
    if name not in _ETYPES:
        raise DeckError(f"element type '{name}' is not supported")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        xyz = np.fromstring(nb, dtype=np.float64, sep=" ")
        el = np.fromstring(eb.replace(b" " + first[j] + b" ", b" "), dtype=np.int64, sep=" ")
    ncol = len(first) - 1
    if xyz.size != 4 * n or el.size != E * ncol:
        return None
    xyz = xyz.reshape(n, 4)
    el = el.reshape(E, ncol)
    if not np.array_equal(xyz[:, 0], np.arange(n)):
        raise DeckError("node ids must be 0..n-1 in order")
    return xyz[:, 1:4].copy(), _ETYPES[name], np.ascontiguousarray(el[:, j:]), el[:, 1].copy()


def read_msh(path: str) -> Tuple[Mesh, np.ndarray]:
    """Nodes, elements and material groups of the first #FEM_MSH."""
    with open(path, "rb") as f:
        fast = _msh_bulk(f.read())
    if fast is not None:
        nodes, etype, elems, groups = fast
        return _make_mesh(nodes, etype, elems), groups
    lines = _data_lines(path)
    try:
        i = next(k for k, s in enumerate(lines) if s.startswith("#FEM_MSH"))
    except StopIteration:
        raise DeckError(f"{path}: no #FEM_MSH")
    nodes = elems = None
    groups = None
    etype = None
    k = i + 1
    while k < len(lines):
        s = lines[k]
        if s.startswith("#FEM_MSH"):
            raise DeckError(f"{path}: several meshes are not supported")
        if s.startswith("$NODES"):
            n = int(lines[k + 1].split()[0])
            body = lines[k + 2:k + 2 + n]
            tok = [b.split()[:4] for b in body] if any("$AREA" in b for b in body) else None
            arr = (np.array(tok, dtype=np.float64) if tok is not None else
                   np.array(" ".join(body).split(), dtype=np.float64).reshape(n, 4))
            if not np.array_equal(arr[:, 0], np.arange(n)):
                raise DeckError(f"{path}: node ids must be 0..n-1 in order")
            nodes = arr[:, 1:4]
            k += 2 + n
            continue
        if s.startswith("$ELEMENTS"):
            n = int(lines[k + 1].split()[0])
            body = [b.split() for b in lines[k + 2:k + 2 + n]]
            et, conn, grp = None, [], []
            for e, t in enumerate(body):
                j = 3 if t[2] == "-1" else 2
                name = t[j]
                if name not in _ETYPES:
                    raise DeckError(f"{path}: element type '{name}' is not supported")
                if et is None:
                    et = _ETYPES[name]
                elif _ETYPES[name] != et:
                    raise DeckError(f"{path}: mixed element types are not supported")
                grp.append(int(t[1]))
                conn.append(t[j + 1:])
            etype = et
            elems = np.array(conn, dtype=np.int64)
            groups = np.array(grp, dtype=np.int64)
            k += 2 + n
            continue
        k += 1
    if nodes is None or elems is None:
        raise DeckError(f"{path}: $NODES and $ELEMENTS are required")
    return _make_mesh(nodes, etype, elems), groups


def _make_mesh(nodes: np.ndarray, etype: int, elems: np.ndarray) -> Mesh:
    if elems.shape[1] != {LINE2: 2, QUAD4: 4, HEX8: 8, TET4: 4}[etype]:
        raise DeckError("element node count does not match its type")
    ed = EDIM[etype]
    x, y, z, flag = _plane_coordinates(nodes, ed)
    mesh = Mesh(etype, ed, x, y, z, elems, (), (), np.arange(elems.shape[0]))
    mesh.coordinate_flag = flag
    mesh.xyz = nodes
    return mesh


def _plane_coordinates(nodes: np.ndarray, edim: int):
    """The coordinate flag of the mesh (10/11/12, 21/22/23, 32 as CFEMesh
    computes it) and the element-local coordinates the assembly uses: the
    axes spanned, in x, y, z order (so the vertical axis, z, comes last)."""
    ext = nodes.max(0) - nodes.min(0)
    on = [bool(ext[a] > 0.0) for a in range(3)]
    axes = [a for a in range(3) if on[a]]
    if len(axes) == 1:
        flag = 10 + axes[0]
    elif len(axes) == 2:
        flag = {(0, 1): 21, (0, 2): 22, (1, 2): 23}[tuple(axes)]
    elif len(axes) == 3:
        flag = 32
    else:
        flag = 10
        axes = [0]
    if len(axes) != edim:
        raise DeckError("elements must span the mesh's dimensions (no lower-dimensional "
                        "elements embedded in 2-D/3-D)")
    c = [nodes[:, a].copy() for a in axes] + [np.zeros(nodes.shape[0])] * (3 - len(axes))
    return c[0], c[1], c[2], flag


@dataclass
class Geometry:
    points: Dict[str, np.ndarray] = field(default_factory=dict)
    point_list: List[np.ndarray] = field(default_factory=list)
    polylines: Dict[str, Tuple[List[int], Optional[float]]] = field(default_factory=dict)


def read_gli(path: str) -> Geometry:
    g = Geometry()
    for kw, lines in _blocks(path):
        if kw == "POINTS":
            for s in lines:
                t = s.split()
                if not t or t[0].startswith("$"):
                    continue
                p = np.array([float(v) for v in t[1:4]])
                g.point_list.append(p)
                if "$NAME" in t:
                    g.points[t[t.index("$NAME") + 1]] = p
        elif kw == "POLYLINE":
            name, ids, eps = None, [], None
            for sk, data in _subkeys(lines):
                if sk == "NAME":
                    name = data[0].split()[0]
                elif sk == "POINTS":
                    ids = [int(d.split()[0]) for d in data]
                elif sk == "EPSILON":
                    eps = float(data[0].split()[0])
            if name is None:
                raise DeckError(f"{path}: polyline without $NAME")
            g.polylines[name] = (ids, eps)
    return g


_EDGES = {
    LINE2: [(0, 1)],
    QUAD4: [(0, 1), (1, 2), (2, 3), (3, 0)],
    HEX8: [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
           (0, 4), (1, 5), (2, 6), (3, 7)],
    TET4: [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)],
}


def _min_edge(mesh: Mesh) -> float:
    """Shortest element edge (CFEMesh::getMinEdgeLength)."""
    X, c = mesh.xyz, mesh.conn
    return float(min(np.linalg.norm(X[c[:, a]] - X[c[:, b]], axis=1).min()
                     for a, b in _EDGES[mesh.etype]))


def geo_nodes(mesh: Mesh, geo: Geometry, gtype: str, name: Optional[str]) -> np.ndarray:
    """Mesh nodes of a geometric object: POINT = the nearest node
    (CFEMesh::GetNODOnPNT, MSH/msh_mesh.cpp:1434); POLYLINE = the nodes within
    the polyline's $EPSILON, else half the shortest element edge, of one of
    its segments (GetNODOnPLY :1685); DOMAIN = every node."""
    X = mesh.xyz
    if gtype == "DOMAIN":
        return np.arange(mesh.n_nodes, dtype=np.int64)
    if gtype == "POINT":
        if name not in geo.points:
            raise DeckError(f"unknown point {name}")
        return np.array([int(np.argmin(((X - geo.points[name]) ** 2).sum(1)))], np.int64)
    if gtype == "POLYLINE":
        if name not in geo.polylines:
            raise DeckError(f"unknown polyline {name}")
        ids, eps = geo.polylines[name]
        r = eps if eps is not None else 0.5 * _min_edge(mesh)
        P = [geo.point_list[i] for i in ids]
        hit = np.zeros(mesh.n_nodes, bool)
        for a, b in zip(P[:-1], P[1:]):
            ab = b - a
            t = np.clip(((X - a) @ ab) / max(ab @ ab, 1e-300), 0.0, 1.0)
            d = np.linalg.norm(X - (a + t[:, None] * ab[None, :]), axis=1)
            hit |= d <= r
        return np.nonzero(hit)[0].astype(np.int64)  # node order
    raise DeckError(f"geometry type {gtype} is not supported")


# ---------------------------------------------------------------------------
# process-side files
# ---------------------------------------------------------------------------
@dataclass
class Numerics:
    method: int = 2
    error_method: int = 1
    tol: float = 1e-10
    max_iter: int = 1000
    theta: float = 1.0
    precond: int = 1
    storage: int = 2
    n_gauss: int = 3
    nls_relaxation: float = 0.0


@dataclass
class Timing:
    start: float = 0.0
    end: float = 0.0
    steps: List[float] = field(default_factory=list)
    unit_factor: float = 1.0


@dataclass
class Condition:
    process: str
    variable: str
    gtype: str
    gname: Optional[str]
    value: float


@dataclass
class OutputSpec:
    process: Optional[str]
    values: List[str]
    gtype: str
    gname: Optional[str]
    dat_type: str
    times: List[float]
    every: Optional[int]


def _per_process(path: str, keyword: str):
    out = []
    for kw, lines in _blocks(path):
        if kw != keyword:
            continue
        subs = _subkeys(lines)
        pcs = next((d[0].split()[0] for k, d in subs if k == "PCS_TYPE" and d), None)
        out.append((pcs, subs))
    return out


def read_num(path: str) -> Dict[str, Numerics]:
    res = {}
    for pcs, subs in _per_process(path, "NUMERICS"):
        n = Numerics()
        for k, d in subs:
            if k == "LINEAR_SOLVER" and d:
                t = d[0].split()
                n.method, n.error_method = int(t[0]), int(t[1])
                n.tol, n.max_iter, n.theta = float(t[2]), int(t[3]), float(t[4])
                n.precond = int(t[5]) if len(t) > 5 else 1
                n.storage = int(t[6]) if len(t) > 6 else 2
            elif k == "ELE_GAUSS_POINTS" and d:
                n.n_gauss = int(d[0].split()[0])
            elif k.startswith("NON_LINEAR_ITERATION") and d:
                t = d[0].split()
                n.nls_relaxation = float(t[3]) if len(t) > 3 else 0.0
        res[pcs] = n
    return res


def read_tim(path: str) -> Dict[str, Timing]:
    res = {}
    for pcs, subs in _per_process(path, "TIME_STEPPING"):
        t = Timing()
        for k, d in subs:
            if k == "TIME_START" and d:
                t.start = float(d[0].split()[0])
            elif k == "TIME_END" and d:
                t.end = float(d[0].split()[0])
            elif k == "TIME_STEPS":
                for s in d:
                    c, dt = s.split()[:2]
                    t.steps += [float(dt)] * int(c)
            elif k == "TIME_UNIT" and d:
                u = d[0].split()[0]
                for name, f in _TIME_UNIT:
                    if name in u:
                        t.unit_factor = f
        res[pcs] = t
    return res


def read_conditions(path: str, keyword: str) -> List[Condition]:
    res = []
    if not os.path.exists(path):
        return res
    for pcs, subs in _per_process(path, keyword):
        var, gtype, gname, value = None, None, None, None
        for k, d in subs:
            if k == "PRIMARY_VARIABLE" and d:
                var = d[0].split()[0]
            elif k == "GEO_TYPE" and d:
                # only the first line names the object (GeoIO::readGeoInfo,
                # FileIO/FEMIO/GeoIO.cpp:22; CInitialCondition::Read,
                # FEM/rf_ic_new.cpp:363): further lines are not keywords and
                # are skipped
                t = d[0].split()
                gtype, gname = t[0], (t[1] if len(t) > 1 else None)
            elif k == "DIS_TYPE" and d:
                t = d[0].split()
                if t[0] != "CONSTANT":
                    raise DeckError(f"{path}: $DIS_TYPE {t[0]} is not supported (CONSTANT only)")
                value = float(t[1])
        if gtype is None or value is None:
            raise DeckError(f"{path}: condition without $GEO_TYPE or $DIS_TYPE")
        res.append(Condition(pcs, var, gtype, gname, value))
    return res


@dataclass
class Medium:
    geo_dimension: int = 0
    geo_area: float = 1.0
    porosity: float = 0.0
    storage: float = 0.0
    perm: Optional[np.ndarray] = None  # dim x dim


def read_mmp(path: str, dim: int) -> List[Medium]:
    res = []
    for kw, lines in _blocks(path):
        if kw != "MEDIUM_PROPERTIES":
            continue
        m = Medium()
        for k, d in _subkeys(lines):
            if k == "GEOMETRY_DIMENSION" and d:
                m.geo_dimension = int(d[0].split()[0])
            elif k == "GEOMETRY_AREA" and d:
                m.geo_area = float(d[0].split()[0])
            elif k == "POROSITY" and d:
                m.porosity = _model_value(d, "$POROSITY")
            elif k == "STORAGE" and d:
                m.storage = _model_value(d, "$STORAGE")
            elif k == "PERMEABILITY_TENSOR" and d:
                t = d[0].split()
                v = [float(s) for s in t[1:]]
                if t[0].startswith("I"):
                    m.perm = np.eye(dim) * v[0]
                elif t[0].startswith("O"):
                    m.perm = np.diag(v[:dim])
                elif t[0].startswith("A"):
                    m.perm = np.array(v[:dim * dim]).reshape(dim, dim)
                else:
                    raise DeckError(f"{path}: permeability tensor type {t[0]}")
        res.append(m)
    return res


@dataclass
class Solid:
    density: float = 0.0
    capacity: float = 0.0
    conductivity: float = 0.0


def read_msp(path: str) -> List[Solid]:
    res = []
    if not os.path.exists(path):
        return res
    for kw, lines in _blocks(path):
        if kw != "SOLID_PROPERTIES":
            continue
        s = Solid()
        for i, line in enumerate(lines):
            nxt = lines[i + 1:i + 2]
            if line.startswith("$DENSITY"):
                s.density = _model_value(nxt, "solid $DENSITY")
            elif "CAPACITY" in line and not line[:1].isdigit():
                s.capacity = _model_value(nxt, "solid CAPACITY")
            elif line == "CONDUCTIVITY":
                s.conductivity = _model_value(nxt, "solid CONDUCTIVITY")
        res.append(s)
    return res


@dataclass
class Fluid:
    density: float = 0.0
    viscosity: float = 1e-3
    heat_capacity: float = 0.0
    heat_conductivity: float = 0.0
    gravity: float = GRAVITY_CONSTANT
    cal_gravity: bool = True


def read_mfp(path: str) -> Fluid:
    f = Fluid()
    if not os.path.exists(path):
        return f
    for kw, lines in _blocks(path):
        if kw != "FLUID_PROPERTIES":
            continue
        for k, d in _subkeys(lines):
            if k == "DENSITY":
                f.density = _model_value(d, "fluid $DENSITY")
            elif k == "VISCOSITY":
                f.viscosity = _model_value(d, "fluid $VISCOSITY")
            elif k == "SPECIFIC_HEAT_CAPACITY":
                f.heat_capacity = _model_value(d, "fluid $SPECIFIC_HEAT_CAPACITY")
            elif k == "HEAT_CONDUCTIVITY":
                f.heat_conductivity = _model_value(d, "fluid $HEAT_CONDUCTIVITY")
            elif k == "GRAVITY" and d:
                f.gravity = float(d[0].split()[0])
            elif k == "NON_GRAVITY":
                f.cal_gravity = False
        break  # the first fluid is the liquid phase
    return f


def read_out(path: str) -> List[OutputSpec]:
    res = []
    if not os.path.exists(path):
        return res
    for kw, lines in _blocks(path):
        if kw != "OUTPUT":
            continue
        pcs, vals, gtype, gname, dat, times, every = None, [], "DOMAIN", None, "TECPLOT", [], None
        for k, d in _subkeys(lines):
            if k == "PCS_TYPE" and d:
                pcs = d[0].split()[0]
            elif k == "NOD_VALUES":
                vals = [s.split()[0] for s in d]
            elif k == "GEO_TYPE" and d:
                t = d[0].split()
                gtype, gname = t[0], (t[1] if len(t) > 1 else None)
            elif k == "DAT_TYPE" and d:
                dat = d[0].split()[0]
            elif k == "TIM_TYPE" and d:
                t = d[0].split()
                if t[0] == "STEPS":
                    every = int(t[1])
                else:
                    times = [float(s.split()[0]) for s in d]
        res.append(OutputSpec(pcs, vals, gtype, gname, dat, times, every))
    return res


# ---------------------------------------------------------------------------
# the deck
# ---------------------------------------------------------------------------
@dataclass
class Deck:
    base: str
    processes: List[str]
    mesh: Mesh
    groups: np.ndarray
    geo: Geometry
    num: Dict[str, Numerics]
    tim: Dict[str, Timing]
    bcs: List[Condition]
    sts: List[Condition]
    ics: List[Condition]
    media: List[Medium]
    solids: List[Solid]
    fluid: Fluid
    outputs: List[OutputSpec]

    @property
    def has_flow(self) -> bool:
        return any(p in _FLOW for p in self.processes)

    def scenario(self, pcs: str) -> Scenario:
        """The process's time loop as a Scenario (element coefficients,
        Dirichlet and source lists, initial condition, steps, solver)."""
        if pcs not in _PROCESS:
            raise DeckError(f"process {pcs} is not supported")
        proc = _PROCESS[pcs]
        m, g = self.mesh, self.groups
        if g.max() >= len(self.media):
            raise DeckError("an element's material group has no #MEDIUM_PROPERTIES")
        num = self.num.get(pcs, self.num.get(None, Numerics()))
        tim = self.tim.get(pcs) or next(iter(self.tim.values()))
        D = m.dim
        fl = self.fluid
        geo_area = self.media[0].geo_area
        if any(md.geo_area != geo_area for md in self.media):
            raise DeckError("one $GEOMETRY_AREA for all material groups is supported")
        if proc in (L.LIQUID_FLOW, L.GROUNDWATER_FLOW):
            mass = np.array([self.media[k].storage for k in g])
            perms = []
            for md in self.media:
                if md.perm is None:
                    raise DeckError(f"{pcs} needs $PERMEABILITY_TENSOR")
                perms.append(md.perm)
            iso = all(np.array_equal(p, np.eye(D) * p[0, 0]) for p in perms)
            cond = (np.array([perms[k][0, 0] for k in g]) if iso else
                    np.stack([perms[k].ravel() for k in g]))
            stride = 1 if iso else D * D
            vertical = m.coordinate_flag % 10 == 2
            rho_g = fl.gravity * fl.density if (vertical and fl.cal_gravity and
                                                fl.gravity >= 1e-20) else 0.0
            theta = num.nls_relaxation
            if proc == L.GROUNDWATER_FLOW:  # HEAD form: K is the conductivity
                extra = dict(mu=1.0, rho_g=0.0)
            else:
                if fl.viscosity == 0.0:
                    raise DeckError("LIQUID_FLOW needs a nonzero $VISCOSITY")
                extra = dict(mu=fl.viscosity, rho_g=rho_g)
        else:
            poro = np.array([self.media[k].porosity for k in g]) if self.has_flow else \
                np.zeros(m.n_elem)
            sol = [self.solids[k] if k < len(self.solids) else None for k in g]
            if any(s is None for s in sol):
                raise DeckError("HEAT_TRANSPORT needs #SOLID_PROPERTIES for every material group")
            rho_c_s = np.array([s.capacity * abs(s.density) for s in sol])
            fluid_rc = fl.density * fl.heat_capacity if self.has_flow else 0.0
            mass = poro * fluid_rc + (1.0 - poro) * rho_c_s
            lam_f = fl.heat_conductivity if self.has_flow else 0.0
            cond = np.array([s.conductivity for s in sol]) * (1.0 - poro) + poro * lam_f
            stride = 1
            theta = num.theta
            extra = dict(adv_coef=fl.heat_capacity * fl.density if self.has_flow else 0.0)
        u0 = np.zeros(m.n_nodes)
        for c in self.ics:
            if c.process == pcs:
                u0[geo_nodes(m, self.geo, c.gtype, c.gname)] = c.value
        bc_idx, bc_val = [], []
        for c in self.bcs:
            if c.process == pcs:
                nodes = geo_nodes(m, self.geo, c.gtype, c.gname)
                bc_idx.append(nodes)
                bc_val.append(np.full(nodes.size, c.value))
        st_idx, st_val = [], []
        for c in self.sts:
            if c.process == pcs:
                if c.gtype != "POINT":
                    raise DeckError("source terms: POINT only")
                nodes = geo_nodes(m, self.geo, c.gtype, c.gname)
                st_idx.append(nodes)
                st_val.append(np.full(nodes.size, c.value))
        cat = lambda a, t: np.concatenate(a).astype(t) if a else np.zeros(0, t)
        sc = Scenario(f"{os.path.basename(self.base)}:{pcs}", m, proc, mass.astype(np.float64),
                      np.ascontiguousarray(cond, dtype=np.float64), stride,
                      cat(bc_idx, np.int64), cat(bc_val, np.float64),
                      cat(st_idx, np.int64), cat(st_val, np.float64), u0,
                      [Step(dt, theta) for dt in tim.steps], tol=num.tol,
                      max_iter=num.max_iter, n_gauss=num.n_gauss, tuf=tim.unit_factor,
                      geo_area=geo_area, **extra)
        sc.meta.update(method=num.method, precond=num.precond, time_start=tim.start,
                       variable=_PRIMARY[proc])
        return sc


def read_deck(base: str) -> Deck:
    """Reads <base>.pcs, .num, .tim, .msh, .gli, .bc, .st, .ic, .mmp, .msp,
    .mfp, .out (the ones present)."""
    f = lambda ext: base + ext
    procs = []
    for pcs, _ in _per_process(f(".pcs"), "PROCESS"):
        if pcs not in _PROCESS:
            raise DeckError(f"process {pcs} is not supported")
        procs.append(pcs)
    if not procs:
        raise DeckError(f"{f('.pcs')}: no process")
    mesh, groups = read_msh(f(".msh"))
    geo = read_gli(f(".gli")) if os.path.exists(f(".gli")) else Geometry()
    return Deck(base, procs, mesh, groups, geo, read_num(f(".num")), read_tim(f(".tim")),
                read_conditions(f(".bc"), "BOUNDARY_CONDITION"),
                read_conditions(f(".st"), "SOURCE_TERM"),
                read_conditions(f(".ic"), "INITIAL_CONDITION"),
                read_mmp(f(".mmp"), mesh.dim), read_msp(f(".msp")), read_mfp(f(".mfp")),
                read_out(f(".out")))


# ---------------------------------------------------------------------------
# running a deck on the device
# ---------------------------------------------------------------------------
def run_deck(base: str, out_dir: Optional[str] = None, precision: int = L.FP64) -> dict:
    """Runs the deck's processes, in .pcs order, through every time step
    (Problem::Euler_TimeDiscretize -> CRFProcess::Execute per process),
    LIQUID_FLOW's Gauss-point velocity advecting HEAT_TRANSPORT, and writes
    the deck's outputs (TECPLOT point curves and domain files, VTK domain
    files) under out_dir (default: next to the deck)."""
    from . import output as O
    deck = read_deck(base)
    scs = [deck.scenario(p) for p in deck.processes]
    pat = L.Pattern(deck.mesh.n_nodes, deck.mesh.etype, deck.mesh.conn)
    systems = []
    for sc in scs:
        s = L.System(pat, precision)
        s.set_elements(**sc.element_kwargs())
        s.set_vector(L.VEC_UOLD, sc.u0)
        s.set_vector(L.VEC_X, sc.u0)
        systems.append(s)
    out_base = os.path.join(out_dir or os.path.dirname(base) or ".", os.path.basename(base))
    writer = O.DeckOutput(deck, out_base)
    values = {sc.meta["variable"]: sc.u0.copy() for sc in scs}
    t = scs[0].meta["time_start"]
    writer.write(0, t, values)
    infos = {p: [] for p in deck.processes}
    n_steps = max(len(sc.steps) for sc in scs)
    flow_i = next((i for i, p in enumerate(deck.processes) if p in _FLOW), None)
    heat_i = deck.processes.index("HEAT_TRANSPORT") if "HEAT_TRANSPORT" in deck.processes else None
    try:
        for k in range(n_steps):
            for i, (pcs, sc, s) in enumerate(zip(deck.processes, scs, systems)):
                if k >= len(sc.steps):
                    continue
                st = sc.steps[k]
                s.initialize()
                s.assemble(st.dt, st.theta, st.steady)
                if sc.st_idx.size:
                    s.add_rhs(sc.st_idx, sc.st_val)
                cg = sc.meta["method"] == L.CG
                s.set_dirichlet(sc.bc_idx, sc.bc_val,
                                L.DIRICHLET_SYMMETRIC if cg else L.DIRICHLET_ROW)
                info = s.solve(method=L.CG if cg else L.BICGSTAB, tol=sc.tol,
                               max_iter=sc.max_iter,
                               precond=L.PRECOND_JACOBI if sc.meta["precond"] in (1, 100)
                               else L.PRECOND_NONE)
                infos[pcs].append(info)
                s.copy_vector(L.VEC_UOLD, L.VEC_X)
                values[sc.meta["variable"]] = s.get_vector(L.VEC_X)
                if i == flow_i and heat_i is not None:
                    s.gauss_velocity(systems[heat_i], host=False)
            t += scs[0].steps[k].dt if k < len(scs[0].steps) else 0.0
            writer.write(k + 1, t, values)
    finally:
        for s in systems:
            s.close()
    return dict(deck=deck, values=values, infos=infos, files=writer.files, time=t)
