"""The serving loop: requests of different sizes arrive while others run,
are admitted by slots and HBM footprint, and share the GPU through one
group.  Each served result must equal the same request solved alone."""
import numpy as np
import pytest

import ogs5_amd as ox
from ogs5_amd import workloads as W
from ogs5_amd.serve import Server

pytestmark = pytest.mark.gpu


def _stream(count=14, seed=3):
    return W.request_stream(count, rate=400.0, seed=seed, quad_sides=(8, 48), hex_sides=(4, 12),
                            steps=(1, 4))


@pytest.mark.parametrize("prec", [ox.FP32_IR, ox.FP64])
def test_served_results_equal_alone(prec):
    specs = _stream()
    # a small budget and few slots force queueing and joins mid-solve
    srv = Server(prec, hbm_budget=64 << 20, max_active=4, batch_iters=8, workers=4)
    out = srv.serve(specs, keep_x=True)
    assert out["requests"] == len(specs) and not out["not_converged"]
    assert out["peak_resident_requests"] <= 4 and out["peak_resident_bytes"] <= 64 << 20
    print(f"[serve] prec={prec} p50={out['latency_ms']['p50']:.1f} ms "
          f"p99={out['latency_ms']['p99']:.1f} ms peak={out['peak_resident_requests']}",
          flush=True)
    for job in out["jobs"]:
        alone = W.run_gpu(job.sc, precision=prec, tol=1e-10)
        assert np.array_equal(alone["x"], job.x), job.spec
        assert [i["iterations"] for i in alone["info"]] == [i["iterations"] for i in job.infos]


def test_footprint_matches_allocation():
    for sc in (W.quad_flow(30, 20, n_steps=1), W.hex_flow(7, n_transient=1)):
        for prec in (ox.FP64, ox.FP32_IR):
            pat, s = W.make_system(sc, prec)
            assert pat.footprint(prec) == s.device_bytes()
