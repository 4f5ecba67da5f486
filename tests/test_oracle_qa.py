"""The oracle against the reference's QA regression logs (QA/TEST_*/REF/stdout.0).

Every per-step row (Time %f, Dt %e, Ek %e, Om %e, ReLambda %8.3f, as printed by
src/MAIN.cpp:440) must match the reference's printed digits exactly.  TEST_3..9
start from the K41 field drawn per rank (src/HIT_LES.cpp:502-540); the oracle
emulates the 4-rank slab split the REF logs were made with.
"""
import pytest

import qa_driver
from oracle.oracle import Oracle

FIX = qa_driver.load_fixtures()


@pytest.mark.parametrize("name", sorted(FIX))
def test_qa_log_rows(name):
    fx = FIX[name]
    rows, _ = qa_driver.run(fx["params"], lambda **kw: Oracle(**kw))
    ref = fx["rows"]
    assert len(rows) == len(ref)
    bad = [(i, r, e) for i, (r, e) in enumerate(zip(rows, ref)) if r != e]
    assert not bad, bad[:3]
    assert fx["final"]["Final time-steps"] == str(len(rows))
