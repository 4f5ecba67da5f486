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
    biggest = max(device_bytes(**r.solver_args(PREC_FP32_SPLIT)) for r in reqs)
    srv = serve.Server(hbm_budget=biggest * 1.01, chunk=5)
    stats, done = srv.run(reqs)
    assert stats["requests"] == 6
    # with room for one run at a time, runs never overlap in time
    spans = sorted((r.admitted, r.finished) for r in done)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1 - 1e-9
