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
    # requests 4-7 take the shape classes of 0-3 and arrive after those have
    # finished, so they run on kept (reconfigured) contexts
    from spns_amd import serve
    reqs = serve.draw_requests(8, rate=200.0, seed=3, steps=(20, 60))
    for i in range(4):
        reqs[4 + i].grid, reqs[4 + i].C_Smag = reqs[i].grid, reqs[i].C_Smag
        reqs[4 + i].arrival = 2.0 + 0.01 * i
    srv = serve.Server(hbm_budget=40e9, chunk=7, batched=batched)
    stats, done = srv.run(reqs, keep_state=True)
    assert stats["requests"] == 8
    assert [r.error for r in done if r.error] == []
    assert stats["contexts_reused"] >= 1  # (all four when the first runs finish before the later ones arrive)
    for r in done:
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


@pytest.mark.parametrize("les", [False, True])
def test_reconfigured_context_matches_fresh(les):
    """A finished run's context reused for another run of the same shape
    class (spns_reconfigure, as the serving loop does) gives bit for bit what
    a fresh context gives; anything but the physics must stay the same."""
    import spns_amd
    base = dict(Nx=32, C_At=0.5, precision=1)
    a = spns_amd.Solver(nu=1e-3, C_Smag=0.17 if les else 0.0, **base)
    a.init_random(seed=3, k0=4.0)
    a.finish_init(False)
    a.step(30)
    phys = dict(nu=3e-3, C_Smag=0.12 if les else 0.0, omega=(0.0, 0.0, 5.0))
    a.reconfigure(**phys)
    b = spns_amd.Solver(**dict(base, **phys))
    for s in (a, b):
        s.init_random(seed=9, k0=4.0)
        s.finish_init(False)
        s.step(25)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.get_rhist(), b.get_rhist())
    assert np.array_equal(a.spectrum(), b.spectrum())
    sa, sb = a.scalars(), b.scalars()
    assert all(sa[k] == sb[k] for k in ("time", "dt", "ek", "om", "iter"))
    with pytest.raises(spns_amd.SpnsError) as e:
        a.reconfigure(Nx=64)
    assert e.value.code == spns_amd.SPNS_EINVAL
    with pytest.raises(spns_amd.SpnsError):
        a.reconfigure(C_Smag=0.0 if les else 0.15)  # DNS <-> LES changes the buffers
    a.close()
    b.close()
