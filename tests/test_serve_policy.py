"""The serving loop's choice of which runs get their next chunk
(spns_amd.serve.eligible_runs), on the host: processor sharing takes every
run; the shortest-remaining-work policy takes runs in ascending order of
step_cost x steps left while their per-step costs fit the budget, and at
least min_inflight of them."""
from spns_amd.serve import Request, eligible_runs, step_cost


def run(rid, grid, steps, done=0, les=False):
    r = Request(rid=rid, seed=rid, grid=grid, C_Smag=0.17 if les else 0.0, nu=1e-3, omega_z=0.0, steps=steps,
                arrival=0.0)
    return {"req": r, "done": done}


def test_step_cost():
    assert step_cost(run(0, (32, 32, 32), 10)["req"]) == 1e6  # launch-bound floor
    big = run(1, (192, 128, 128), 10)["req"]
    assert step_cost(big) == 288 * 192 * 192
    assert step_cost(run(2, (192, 128, 128), 10, les=True)["req"]) == 1.4 * 288 * 192 * 192


def test_processor_sharing_takes_all():
    act = [run(i, (64, 64, 64), 100) for i in range(5)]
    assert eligible_runs(act, None, 1) == set(range(5))


def test_shortest_remaining_work_first():
    act = [run(0, (192, 128, 128), 2000), run(1, (32, 32, 32), 500, done=400), run(2, (32, 32, 32), 1000),
           run(3, (128, 128, 128), 100), run(4, (64, 64, 64), 300)]
    # remaining work: 0: 10.6e6*2000, 1: 1e6*100, 2: 1e6*1000, 3: 7.08e6*100, 4: 1e6*300
    # order 1, 4, 3, 2, 0; per-step costs 1e6, 1e6, 7.08e6, 1e6, 10.6e6
    assert eligible_runs(act, 2e6, 1) == {1, 4}
    assert eligible_runs(act, 1e7, 1) == {1, 4, 3}      # run 2 would exceed: the prefix stops there
    assert eligible_runs(act, 1.1e7, 1) == {1, 4, 3, 2}
    assert eligible_runs(act, 1e6, 3) == {1, 4, 3}      # min_inflight overrides the budget
    assert eligible_runs(act, 1e12, 1) == set(range(5))
    assert eligible_runs(act, 0.0, 1) == {1}            # at least one run always advances


def test_finished_runs_sort_first():
    act = [run(0, (64, 64, 64), 100, done=100), run(1, (32, 32, 32), 100)]
    assert 0 in eligible_runs(act, 1.0, 1)  # no remaining work: retired on its turn


def test_long_run_is_not_starved():
    """A long run among a steady stream of short arrivals: with aging it gets
    a chunk at least once every max_wait + 1 rounds; without aging it would
    wait as long as short runs keep arriving."""
    long = run(0, (192, 128, 128), 2000)
    long["waited"] = 0
    for max_wait, expect_starved in ((4, False), (None, True)):
        act = [dict(long, done=0, waited=0)]
        nid, gaps, last = 1, [], 0
        for rnd in range(1, 60):
            # two short runs arrive every round and need a few rounds each
            for _ in range(2):
                a = run(nid, (32, 32, 32), 40)
                a["waited"] = 0
                act.append(a)
                nid += 1
            chosen = eligible_runs(act, 4e6, 2, max_wait)
            for i, a in enumerate(act):
                if i in chosen:
                    a["done"] += 20
                    a["waited"] = 0
                else:
                    a["waited"] = a.get("waited", 0) + 1
            if 0 in chosen and act[0]["req"].rid == 0:
                gaps.append(rnd - last)
                last = rnd
            act = [a for a in act if a["done"] < a["req"].steps or a["req"].rid == 0]
        if expect_starved:
            assert not gaps  # shortest-remaining-work alone never picks the long run here
        else:
            assert gaps and max(gaps) <= max_wait + 1
            assert act[0]["done"] >= 20 * (59 // (max_wait + 1))


def test_overdue_runs_do_not_count_against_min_inflight():
    act = [run(0, (192, 128, 128), 2000), run(1, (32, 32, 32), 100), run(2, (32, 32, 32), 100)]
    for a in act:
        a["waited"] = 0
    act[0]["waited"] = 3
    assert eligible_runs(act, 1e6, 1, max_wait=3) == {0, 1}
    assert eligible_runs(act, 1e6, 1, max_wait=4) == {1}
