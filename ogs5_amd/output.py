"""Output of node values in the reference's file formats.

* TECPLOT time curves in points (COutput::NODWritePNTDataTEC,
  FEM/Output.cpp:2352): `<base>_time_<POINT>[_<PROCESS>].tec`, header at
  step 0 (TITLE, VARIABLES "TIME" and the values, ZONE T="POINT=<name>"),
  then one line per time step: time and the node's values, scientific with
  12 digits.
* TECPLOT domain files (NODWriteDOMDataTEC :858, WriteTECHeader :1439):
  `<base>_domain[_<PROCESS>]_<line|quad|hex|tet>.tec`, per output time a
  zone: VARIABLES "X","Y","Z" and the values, ZONE T="<t>s", N, E,
  F=FEPOINT, ET=QUADRILATERAL|BRICK|TETRAHEDRON, STRANDID/SOLUTIONTIME, the
  node lines and the 1-based connectivity (line elements written as
  degenerate quadrilaterals n0 n1 n1 n0).
* Legacy VTK (LegacyVtkInterface::WriteDataVTK,
  FileIO/MeshIO/LegacyVtkInterface.cpp:709): `<base>[_<PROCESS>]<step:04d>.vtk`,
  UNSTRUCTURED_GRID with the TimesAndCycles field, POINTS, CELLS,
  CELL_TYPES (3 line, 9 quad, 12 hex, 10 tet) and POINT_DATA scalars.
"""
from __future__ import annotations

import os
from typing import Dict, List

import numpy as np

from .mesh import HEX8, LINE2, QUAD4, TET4

_TEC_ET = {LINE2: ("line", "QUADRILATERAL"), QUAD4: ("quad", "QUADRILATERAL"),
           HEX8: ("hex", "BRICK"), TET4: ("tet", "TETRAHEDRON")}
_VTK_CELL = {LINE2: 3, QUAD4: 9, HEX8: 12, TET4: 10}


def _e(v: float) -> str:
    return f"{v:.12e}"


def point_curve_header(name: str, values: List[str]) -> str:
    return (' TITLE = "Time curves in points"\n VARIABLES = "TIME" ' +
            "".join(f' "{v}" ' for v in values) + f'\n ZONE T="POINT={name}"\n')


def point_curve_line(t: float, vals: List[float]) -> str:
    return _e(t) + " " + "".join(_e(v) + " " for v in vals) + "\n"


def tecplot_domain_zone(mesh, xyz: np.ndarray, names: List[str], cols: List[np.ndarray],
                        t: float) -> str:
    et = _TEC_ET[mesh.etype][1]
    out = ['VARIABLES  = "X","Y","Z"' + "".join(f', "{n}"' for n in names),
           f'ZONE T="{t:g}s", N={mesh.n_nodes}, E={mesh.n_elem}, F=FEPOINT, ET={et}',
           f"STRANDID=1, SOLUTIONTIME={t:g}"]
    data = np.column_stack([xyz] + [np.asarray(c, dtype=np.float64) for c in cols])
    out += [" ".join(_e(v) for v in row) + " " for row in data]
    conn = mesh.conn + 1
    if mesh.etype == LINE2:
        conn = conn[:, [0, 1, 1, 0]]
    out += [" ".join(str(int(v)) for v in row) + " " for row in conn]
    return "\n".join(out) + "\n"


def vtk_legacy(mesh, xyz: np.ndarray, names: List[str], cols: List[np.ndarray], t: float,
               step: int) -> str:
    n, E, npe = mesh.n_nodes, mesh.n_elem, mesh.conn.shape[1]
    out = ["# vtk DataFile Version 3.0", "Unstructured Grid from OpenGeoSys", "ASCII",
           "DATASET UNSTRUCTURED_GRID", "FIELD TimesAndCycles 2", "TIME 1 1 double", _e(t),
           "CYLCE 1 1 long", str(step), f"POINTS {n} double"]
    out += [f"{_e(p[0])} {_e(p[1])} {_e(p[2])}" for p in xyz]
    out.append(f"CELLS {E} {E * (npe + 1)}")
    out += [f"{npe} " + " ".join(str(int(v)) for v in row) for row in mesh.conn]
    out.append("")
    out.append(f"CELL_TYPES {E}")
    out += [str(_VTK_CELL[mesh.etype])] * E
    out.append(f"POINT_DATA {n}")
    for name, c in zip(names, cols):
        out += [f"SCALARS {name} double 1", "LOOKUP_TABLE default"]
        out += [_e(v) for v in np.asarray(c, dtype=np.float64)]
    return "\n".join(out) + "\n"


class DeckOutput:
    """Writes a deck's #OUTPUT requests as its time loop advances."""

    def __init__(self, deck, out_base: str):
        self.deck = deck
        self.base = out_base
        self.files: List[str] = []
        self.xyz = deck.mesh.xyz

    def _due(self, spec, step: int, t: float) -> bool:
        if spec.gtype == "POINT":
            return True  # breakthrough curves: every step
        if spec.every is not None:
            return step % spec.every == 0
        return any(abs(t - tt) <= 1e-12 * max(1.0, abs(tt)) for tt in spec.times)

    def write(self, step: int, t: float, values: Dict[str, np.ndarray]):
        from .deck import geo_nodes
        for i, spec in enumerate(self.deck.outputs):
            if not self._due(spec, step, t):
                continue
            names = [v for v in spec.values if v in values]
            cols = [values[v] for v in names]
            suffix = f"_{spec.process}" if spec.process else ""
            if spec.gtype == "POINT":
                node = int(geo_nodes(self.deck.mesh, self.deck.geo, "POINT", spec.gname)[0])
                path = f"{self.base}_time_{spec.gname}{suffix}.tec"
                mode = "w" if step == 0 else "a"
                with open(path, mode) as f:
                    if step == 0:
                        f.write(point_curve_header(spec.gname, names))
                    f.write(point_curve_line(t, [c[node] for c in cols]))
            elif spec.gtype == "DOMAIN" and spec.dat_type == "TECPLOT":
                path = f"{self.base}_domain{suffix}_{_TEC_ET[self.deck.mesh.etype][0]}.tec"
                with open(path, "w" if path not in self.files else "a") as f:
                    f.write(tecplot_domain_zone(self.deck.mesh, self.xyz, names, cols, t))
            elif spec.gtype == "DOMAIN" and spec.dat_type in ("VTK", "PVD"):
                path = f"{self.base}{suffix}{step:04d}.vtk"
                with open(path, "w") as f:
                    f.write(vtk_legacy(self.deck.mesh, self.xyz, names, cols, t, step))
            else:
                raise ValueError(f"output {spec.gtype}/{spec.dat_type} is not supported")
            if path not in self.files:
                self.files.append(path)
