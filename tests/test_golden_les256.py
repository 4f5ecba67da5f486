"""Structure of the committed 256^3 LES oracle fixture (written on the host by
tests/golden/make_les256_fixture.py): what the GPU test reads is there, with
the shapes and the invariants the run fixes."""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from test_gpu_les256 import FIXTURE, checkpoints  # noqa: E402


@pytest.fixture(scope="module")
def fx():
    if not os.path.exists(FIXTURE):
        pytest.skip("fixture not written")
    return np.load(FIXTURE, allow_pickle=False)


def test_params_and_rows(fx):
    nu, cs, cat, delta, k0, seed, steps, N = fx["params"]
    assert (nu, cs, cat, k0, seed, steps, N) == (1e-4, 0.17, 0.5, 6.0, 6.0, 500.0, 256.0)
    assert delta == 2.0 * math.pi / 256
    steps = int(steps)
    for k in ("time", "dt", "ek", "om"):
        assert fx[k].shape == (steps + 1,)
    for k in ("s2", "s3"):
        assert fx[k].shape == (steps,)
    assert checkpoints(fx) == [1, 10, 50, 100, 200, 300, 400, 500]


def test_invariants_of_the_run(fx):
    t, dt = fx["time"], fx["dt"]
    assert t[0] == 0.0
    # each step advances the time by the step's own CFL dt (recorded after it)
    np.testing.assert_allclose(np.diff(t), dt[1:], rtol=1e-12)
    # u_rms = 1 sets the initial energy to 1.5
    assert abs(fx["ek"][0] - 1.5) < 1e-12
    # decaying LES: energy falls every step, the gradient moments stay positive
    assert np.all(np.diff(fx["ek"]) < 0)
    assert np.all(fx["s2"] > 0)
    P = 384.0 ** 3
    sk = (fx["s3"] / P) / (fx["s2"] / P) ** 1.5
    # velocity-derivative skewness: about 0 for the Gaussian initial field,
    # developing towards the -0.4 .. -0.5 of developed turbulence
    assert abs(sk[0]) < 0.05
    assert -0.6 < sk[-1] < -0.3
    assert np.all(np.diff(sk[:100]) < 0)


def test_mode_sets_and_checkpoint_data(fx):
    from make_les256_fixture import low_index
    nk = (257, 257, 129)
    assert np.array_equal(fx["low_idx"], low_index(nk, 16))
    assert np.array_equal(fx["low8_idx"], low_index(nk, 8))
    n8, n16 = len(fx["low8_idx"]), len(fx["low_idx"])
    for n in checkpoints(fx):
        assert fx["low8_%d" % n].shape == (3, n8)
        assert fx["norm_%d" % n].shape == (3,)
        eb = fx["ek_bins_%d" % n]
        # E(K) summed over the shells is the energy after the step
        assert abs(eb.sum() - fx["ek"][n]) <= 1e-12 * fx["ek"][n]
    assert fx["low_500"].shape == (3, n16)
    # the |k| <= 8 modes are a subset of the |k| <= 16 ones, with the same values
    pos = np.searchsorted(fx["low_idx"], fx["low8_idx"])
    assert np.array_equal(fx["low_idx"][pos], fx["low8_idx"])
    assert np.array_equal(fx["low_500"][:, pos], fx["low8_500"])
