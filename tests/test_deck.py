"""The OGS-5 deck reader and the output writers (host side, no GPU).

* The reference's own bmskel deck (tests/golden/bmskel, copied from
  tests/data/bmskel) reads into the scenario the golden test runs.
* A written x-z section deck (tests/deckgen.py) exercises the rest of the
  subset: two processes, two material groups, polylines, a point source,
  time unit DAY, gravity in a vertical section, three outputs; each parsed
  quantity is compared with the value the reference's formulas give for
  the deck's numbers.
* The TECPLOT and VTK writers' layout on a two-element mesh.
"""
import os

import numpy as np
import pytest

from ogs5_amd import deck as D
from ogs5_amd import output as O
from ogs5_amd import workloads as W
from ogs5_amd.mesh import QUAD4, quad_mesh

from deckgen import LX, LZ, NX, NZ, write_section_deck

BMSKEL = os.path.join(os.path.dirname(__file__), "golden", "bmskel", "a")


def test_bmskel_deck_reads_as_the_golden_scenario():
    d = D.read_deck(BMSKEL)
    assert d.processes == ["HEAT_TRANSPORT"]
    sc = d.scenario("HEAT_TRANSPORT")
    ref = W.bmskel()
    for k in ("mass_coef", "cond", "bc_idx", "bc_val", "u0", "st_idx"):
        assert np.array_equal(getattr(sc, k), getattr(ref, k)), k
    assert sc.steps == ref.steps and sc.tol == ref.tol and sc.max_iter == ref.max_iter
    assert sc.meta["method"] == 2 and sc.meta["precond"] == 100
    assert d.mesh.coordinate_flag == 10 and sc.rho_g == 0.0 and sc.tuf == 1.0
    (out,) = d.outputs
    assert (out.gtype, out.gname, out.dat_type, out.values) == \
        ("POINT", "POINT0", "TECPLOT", ["TEMPERATURE1"])


def test_section_deck(tmp_path):
    d = D.read_deck(write_section_deck(str(tmp_path)))
    m = d.mesh
    assert d.processes == ["LIQUID_FLOW", "HEAT_TRANSPORT"]
    assert m.coordinate_flag == 22 and m.etype == QUAD4 and m.dim == 2
    assert np.array_equal(m.y, d.mesh.xyz[:, 2])  # z is the second local axis
    flow, heat = d.scenario("LIQUID_FLOW"), d.scenario("HEAT_TRANSPORT")
    grp = d.groups
    left = grp == 0
    assert left.sum() == NX * NZ // 2
    # LIQUID_FLOW: storage, per-group tensors (one orthotropic group -> full
    # tensors for all), viscosity, gravity on the vertical axis, days
    assert np.array_equal(flow.mass_coef, np.where(left, 1e-9, 2e-9))
    assert flow.cond_stride == 4
    K = flow.cond.reshape(-1, 2, 2)
    assert np.array_equal(K[left], np.tile(np.diag([1e-13, 2e-14]), (left.sum(), 1, 1)))
    assert np.array_equal(K[~left], np.tile(np.eye(2) * 5e-14, ((~left).sum(), 1, 1)))
    assert flow.mu == 1e-3 and flow.rho_g == 9.81 * 1000.0 and flow.tuf == 86400.0
    assert flow.meta["method"] == 5 and flow.tol == 1e-13 and flow.steps == [W.Step(1.0, 0.0)] * 3
    # HEAT_TRANSPORT with a flow process: porosity-weighted capacity and
    # conductivity, advection c_f rho_f, theta = ls_theta
    n = np.where(left, 0.2, 0.3)
    rcs = np.where(left, 2500.0 * 800.0, 2600.0 * 900.0)
    assert np.array_equal(heat.mass_coef, n * (1000.0 * 4200.0) + (1.0 - n) * rcs)
    lam = np.where(left, 2.5, 3.0)
    assert np.array_equal(heat.cond, lam * (1.0 - n) + n * 0.6)
    assert heat.adv_coef == 4200.0 * 1000.0 and heat.steps == [W.Step(1.0, 0.5)] * 3
    assert heat.meta["method"] == 2
    # conditions: polylines on the x = 0 / x = LX edges, the well's node
    X = d.mesh.xyz
    on_left = np.nonzero(X[:, 0] == 0.0)[0]
    on_right = np.nonzero(X[:, 0] == LX)[0]
    assert np.array_equal(np.sort(flow.bc_idx), np.sort(np.r_[on_left, on_right]))
    assert np.array_equal(flow.bc_val[np.isin(flow.bc_idx, on_left)], np.full(NZ + 1, 2e5))
    assert np.array_equal(np.sort(heat.bc_idx), on_left)
    well = int(np.argmin(((X - [40.0, 0.0, 10.0]) ** 2).sum(1)))
    assert list(flow.st_idx) == [well] and list(flow.st_val) == [-1e-6]
    assert np.all(flow.u0 == 1e5) and np.all(heat.u0 == 10.0)
    kinds = [(o.gtype, o.dat_type, o.every, o.times) for o in d.outputs]
    assert kinds == [("POINT", "TECPLOT", None, []), ("DOMAIN", "TECPLOT", 1, []),
                     ("DOMAIN", "VTK", None, [3.0])]


def test_groundwater_section_deck(tmp_path):
    d = D.read_deck(write_section_deck(str(tmp_path), flow="GROUNDWATER_FLOW"))
    assert d.processes == ["GROUNDWATER_FLOW", "HEAT_TRANSPORT"] and d.has_flow
    gw, heat = d.scenario("GROUNDWATER_FLOW"), d.scenario("HEAT_TRANSPORT")
    assert gw.process == 2 and gw.mu == 1.0 and gw.rho_g == 0.0  # HEAD form
    assert gw.meta["variable"] == "HEAD" and gw.tuf == 86400.0
    K = gw.cond.reshape(-1, 2, 2)
    assert np.array_equal(K[d.groups == 0][0], np.diag([1e-6, 2e-7]))
    assert np.all(gw.u0 == 10.0) and set(gw.bc_val) == {20.0, 10.0}
    n = np.where(d.groups == 0, 0.2, 0.3)
    assert np.array_equal(heat.cond, np.where(d.groups == 0, 2.5, 3.0) * (1.0 - n) + n * 0.6)


def test_unsupported_keywords_are_refused(tmp_path):
    base = write_section_deck(str(tmp_path))
    with open(base + ".mmp") as f:
        text = f.read()
    with open(base + ".mmp", "w") as f:
        f.write(text.replace("$STORAGE\n  1 1.0e-9", "$STORAGE\n  2 1 2 3 4"))
    with pytest.raises(D.DeckError):
        D.read_deck(base)
    base2 = write_section_deck(str(tmp_path / "b"))
    with open(base2 + ".bc") as f:
        text = f.read()
    with open(base2 + ".bc", "w") as f:
        f.write(text.replace("CONSTANT 60", "LINEAR 1 60 2 70"))
    with pytest.raises(D.DeckError):
        D.read_deck(base2)


def test_tecplot_and_vtk_layout():
    m = quad_mesh(2, 1, 2.0, 1.0)
    xyz = np.stack([m.x, m.y, m.z], 1)
    T = np.arange(m.n_nodes, dtype=float)
    z = O.tecplot_domain_zone(m, xyz, ["TEMPERATURE1"], [T], 2.5).splitlines()
    assert z[0] == 'VARIABLES  = "X","Y","Z", "TEMPERATURE1"'
    assert z[1] == 'ZONE T="2.5s", N=6, E=2, F=FEPOINT, ET=QUADRILATERAL'
    assert z[2] == "STRANDID=1, SOLUTIONTIME=2.5"
    assert len(z) == 3 + 6 + 2
    assert [float(v) for v in z[4].split()] == [1.0, 0.0, 0.0, 1.0]
    assert [int(v) for v in z[-2].split()] == list(m.conn[0] + 1)
    v = O.vtk_legacy(m, xyz, ["TEMPERATURE1"], [T], 2.5, 7).splitlines()
    assert v[:5] == ["# vtk DataFile Version 3.0", "Unstructured Grid from OpenGeoSys", "ASCII",
                     "DATASET UNSTRUCTURED_GRID", "FIELD TimesAndCycles 2"]
    assert v[8] == "7" and v[9] == "POINTS 6 double"
    i = v.index("CELLS 2 10")
    assert v[i + 1] == "4 " + " ".join(str(c) for c in m.conn[0])
    j = v.index("CELL_TYPES 2")
    assert v[j + 1:j + 3] == ["9", "9"]
    k = v.index("POINT_DATA 6")
    assert v[k + 1:k + 3] == ["SCALARS TEMPERATURE1 double 1", "LOOKUP_TABLE default"]
    assert [float(s) for s in v[k + 3:k + 9]] == list(T)
    h = O.point_curve_header("OBS", ["PRESSURE1"]) + O.point_curve_line(1.0, [2.0])
    assert h.splitlines() == [' TITLE = "Time curves in points"',
                              ' VARIABLES = "TIME"  "PRESSURE1" ',
                              ' ZONE T="POINT=OBS"', "1.000000000000e+00 2.000000000000e+00 "]
