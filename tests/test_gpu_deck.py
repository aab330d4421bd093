"""OGS-5 decks run end to end on the device (ogs5_amd.deck.run_deck): the
reference's bmskel deck reaches its golden value, and the x-z section deck
(LIQUID_FLOW -> Gauss-point velocity -> HEAT_TRANSPORT, tests/deckgen.py)
matches the oracle's coupled chain on the scenarios read from the same
files (1e-10 relative L2), with its three outputs written."""
import os
import shutil

import numpy as np
import pytest

from ogs5_amd import deck as D
from oracle import driver

from deckgen import write_section_deck

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "bmskel")


def test_bmskel_deck_end_to_end(tmp_path):
    for f in os.listdir(GOLDEN):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / f)
    res = D.run_deck(str(tmp_path / "a"))
    path = tmp_path / "a_time_POINT0_HEAT_TRANSPORT.tec"
    assert res["files"] == [str(path)]
    lines = path.read_text().splitlines()
    print(f"[deck] bmskel {lines}", flush=True)
    assert lines[:3] == [' TITLE = "Time curves in points"', ' VARIABLES = "TIME"  "TEMPERATURE1" ',
                         ' ZONE T="POINT=POINT0"']
    rows = [[float(v) for v in ln.split()] for ln in lines[3:]]
    assert rows[0] == [0.0, 273.0]
    assert rows[1][0] == 1.0 and abs(rows[1][1] - 274.0) < 1e-9  # the golden T(POINT0)
    assert res["infos"]["HEAT_TRANSPORT"][0]["converged"] == 1


@pytest.mark.parametrize("flow_pcs,var", [("LIQUID_FLOW", "PRESSURE1"),
                                          ("GROUNDWATER_FLOW", "HEAD")])
def test_section_deck_matches_oracle(tmp_path, flow_pcs, var):
    from oracle import ffi
    base = write_section_deck(str(tmp_path), flow=flow_pcs)
    res = D.run_deck(base)
    d = res["deck"]
    flow, heat = d.scenario(flow_pcs), d.scenario("HEAT_TRANSPORT")
    ref = driver.run_oracle_coupled(flow, heat, ffi, tol=1e-14)
    p, T = res["values"][var], res["values"]["TEMPERATURE1"]
    print(f"[deck] section p in [{p.min():.4g}, {p.max():.4g}], T in [{T.min():.4g}, "
          f"{T.max():.4g}]", flush=True)
    for got, want in ((p, ref["x_flow"]), (T, ref["x_heat"])):
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-10
    assert all(i["converged"] == 1 for v in res["infos"].values() for i in v)
    assert T.max() > 10.5  # the inlet heat moved in
    names = sorted(os.path.basename(f) for f in res["files"])
    assert names == [f"sec_{flow_pcs}0003.vtk", "sec_domain_HEAT_TRANSPORT_quad.tec",
                     "sec_time_OBS.tec"]
    curve = (tmp_path / "sec_time_OBS.tec").read_text().splitlines()
    assert len(curve) == 3 + 4  # header, t = 0, 1, 2, 3
    obs = int(np.argmin(((d.mesh.xyz - [61.0, 0.0, 9.0]) ** 2).sum(1)))
    last = [float(v) for v in curve[-1].split()]
    assert last[0] == 3.0 and abs(last[1] - p[obs]) <= 1e-12 * abs(p[obs])
    dom = (tmp_path / "sec_domain_HEAT_TRANSPORT_quad.tec").read_text()
    assert dom.count("ZONE T=") == 4
