"""Serving loop: continuous batching of mixed runs, batched (one launch per
pass over the runs of a shape class) and per-run; each result is the same
alone and in the mix (bit for bit: runs share no state and the kernels are
deterministic), a diverging run is retired alone, and admission respects the
HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batched", [True, False])
def test_mix_equals_alone(batched):
    from spns_amd import serve
    reqs = serve.draw_requests(8, rate=200.0, seed=3, steps=(20, 60))
    srv = serve.Server(hbm_budget=40e9, chunk=7, batched=batched)
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


@pytest.mark.parametrize("batched", [True, False])
def test_diverging_run_is_isolated(batched):
    """One request with a fixed time step far above its CFL limit overflows
    to Inf/NaN within a few steps.  The loop retires it with an error
    (SPNS_ENONFINITE from spns_query) after that chunk and the other runs
    finish, bit-identical to running alone."""
    from spns_amd import SPNS_ENONFINITE, serve
    reqs = serve.draw_requests(6, rate=500.0, seed=7, steps=(30, 50))
    bad = reqs[2]
    bad.grid, bad.C_Smag, bad.nu, bad.steps = (32, 32, 32), 0.0, 1e-4, 400
    bad.fixed_dt = 5.0  # the CFL limit here is about 1e-2
    srv = serve.Server(hbm_budget=40e9, chunk=10, batched=batched)
    stats, done = srv.run(reqs, keep_state=True)
    assert stats["requests"] == 6 and stats["failed"] == 1
    by = {r.rid: r for r in done}
    assert by[bad.rid].error and by[bad.rid].result["error_code"] == SPNS_ENONFINITE
    assert by[bad.rid].result["steps"] < bad.steps  # stopped early, not after all its steps
    for r in done:
        if r.rid == bad.rid:
            continue
        assert not r.error
        alone = serve.run_alone(r)
        assert np.array_equal(alone["state"], r.result["state"]), r.rid
        assert np.array_equal(alone["ek"], r.result["ek"])


def test_device_memory_query():
    from spns_amd import device_memory
    free, total = device_memory(0)
    assert 0 < free <= total
    assert total > 200e9  # MI355X: 288 GB of HBM3E
