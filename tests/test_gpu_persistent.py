"""Persistent Krylov launches (one cooperative launch per batch of
iterations, grid barriers between phases) against separate launches: the
same per-system chunks and partial sums, so iterates, iteration counts and
residuals must be bit-identical.  The barrier's host-bounded wait is
exercised by a diagnostic whose blocks wait for a barrier that cannot
complete, with the host's give-up flag already set: the launch must end
normally and report that the blocks gave up."""
import numpy as np
import pytest

import ogs5_amd as ox
from ogs5_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture
def persistent_mode():
    def use(mode):
        ox.set_persistent(mode)
    yield use
    ox.set_persistent(0)


def test_cooperative_launch_supported():
    mode, supported = ox.get_persistent()
    assert mode == 0 and supported  # off by default


def test_barrier_gives_up_when_the_host_asks():
    assert ox.persistent_giveup_probe() == 1
    # the device is usable afterwards
    res = W.run_gpu(W.quad_flow(8, 6, seed=2, n_steps=1))
    assert res["info"][0]["converged"] == 1


def _same(a, b):
    assert [i["iterations"] for i in a["info"]] == [i["iterations"] for i in b["info"]]
    assert [i["error"] for i in a["info"]] == [i["error"] for i in b["info"]]
    assert [i["converged"] for i in a["info"]] == [i["converged"] for i in b["info"]]
    for xa, xb in zip(a["steps"], b["steps"]):
        assert np.array_equal(xa, xb)


CASES = [
    ("quad", ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP64, None),
    ("quad", ox.BICGSTAB, ox.DIRICHLET_ROW, ox.FP64, None),
    ("quad", ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP32_IR, ox.IR_RELIABLE),
    ("quad", ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP32_IR, ox.IR_RELIABLE_SR),
    ("hex", ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP64, None),
    ("tet", ox.BICGSTAB, ox.DIRICHLET_ROW, ox.FP64, None),
    ("tet", ox.BICGSTAB, ox.DIRICHLET_ROW, ox.FP32_IR, None),
]


def _scenario(name):
    if name == "quad":
        return W.quad_flow(60, 48, seed=31, n_steps=3)
    if name == "hex":
        return W.hex_flow(14, seed=32, n_transient=2)
    return W.tet_heat(9, seed=33, n_steps=3)


@pytest.mark.parametrize("name,method,bc,prec,ir_mode", CASES)
def test_persistent_equals_separate_launches(persistent_mode, name, method, bc, prec, ir_mode):
    sc = _scenario(name)
    kw = dict(method=method, dirichlet=bc, precision=prec, ir_mode=ir_mode, keep_steps=True)
    persistent_mode(0)
    sep = W.run_gpu(sc, **kw)
    persistent_mode(2)
    per = W.run_gpu(sc, **kw)
    print(f"[persistent] {name} m={method} prec={prec} ir={ir_mode} "
          f"iters={[i['iterations'] for i in per['info']]}", flush=True)
    _same(sep, per)
    assert all(i["converged"] == 1 for i in per["info"])


@pytest.mark.parametrize("prec", [ox.FP64, ox.FP32_IR])
def test_persistent_group_with_joins(persistent_mode, prec):
    """A group whose members join one at a time: the persistent batches
    cover every running member's chunks; results equal separate launches."""
    scs = [W.quad_flow(n, n, seed=40 + n, sigma_ln_k=1.5, n_steps=2) for n in (10, 24, 40, 17)]

    def systems():  # one SpMV format for all, so the batches can be persistent
        out = [W.make_system(sc, prec) for sc in scs]
        for _, s in out:
            s.set_spmv_variant(5)
        return out

    persistent_mode(0)
    sep = W.run_group(scs, precision=prec, keep_steps=True, stagger=5, systems=systems())
    persistent_mode(2)
    per = W.run_group(scs, precision=prec, keep_steps=True, stagger=5, systems=systems())
    for k in range(len(scs)):
        _same(dict(info=sep["info"][k], steps=sep["steps"][k]),
              dict(info=per["info"][k], steps=per["steps"][k]))
