"""Serving loop: continuous batching of mixed runs; each result is the same
alone and in the mix (bit for bit: runs share no state and the kernels are
deterministic), and admission respects the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mix_equals_alone():
    from spns_amd import serve
    reqs = serve.draw_requests(8, rate=200.0, seed=3, steps=(20, 60))
    srv = serve.Server(hbm_budget=40e9, chunk=7)
    stats, done = srv.run(reqs, keep_state=True)
    assert stats["requests"] == 8
    for r in done[:4]:
        alone = serve.run_alone(r)
        assert np.array_equal(alone["state"], r.result["state"]), r.rid
        assert np.array_equal(alone["ek"], r.result["ek"])
        assert alone["time"] == r.result["time"]


def test_admission_respects_budget():
    from spns_amd import serve, device_bytes, PREC_FP32_SPLIT
    reqs = serve.draw_requests(6, rate=1000.0, seed=5, steps=(10, 20))
    fp = {r.rid: device_bytes(**r.solver_args(PREC_FP32_SPLIT)) for r in reqs}
    budget = max(fp.values()) * 1.01
    srv = serve.Server(hbm_budget=budget, chunk=5)
    stats, done = srv.run(reqs)
    assert stats["requests"] == 6
    # at every admission, the runs resident at that moment fit the budget
    for r in done:
        resident = [q for q in done if q.admitted <= r.admitted < q.finished]
        assert sum(fp[q.rid] for q in resident) <= budget
    assert stats["latency_p99_s"] >= stats["latency_p50_s"] > 0
