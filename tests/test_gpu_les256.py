"""The 256^3 Smagorinsky LES at full size, FP64 plain form against the FP64
oracle: k0 = 6 seeded field (u_rms = 1), nu = 1e-4, C_s = 0.17,
Delta = 2 pi / 256, CFL step with C_At = 0.5.  First a few steps stepped
beside the oracle (state, R history, E(k), time, dt, energy and enstrophy
within relative L2 1e-12 after each chunk, skewness too); then the whole
500-step horizon against the oracle's committed run of the same field."""
import math
import os

import numpy as np
import pytest

from test_gpu_parity import compare, pair

pytestmark = pytest.mark.gpu


def test_les_256_fp64_vs_oracle():
    kw = dict(Nx=256, nu=1e-4, C_Smag=0.17, C_At=0.5, delta=2.0 * math.pi / 256)
    g, o = pair(seed=6, k0=6.0, **kw)
    g.finish_init(False)
    o.finish_init(False)
    compare(g, o, "les256 init")
    for chunk in (1, 2):
        g.step(chunk)
        o.step(chunk)
        compare(g, o, "les256 step+%d" % chunk)
    sg, so = g.scalars(), o.scalars()
    P = np.prod(o.shape)
    sk_o = (so["skew_s3"] / P) / (so["skew_s2"] / P) ** 1.5
    assert abs(sg["skewness"] - sk_o) <= 1e-12 * max(1.0, abs(sk_o))
    g.close()


FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "les256_k0_6_seed6.npz")


def checkpoints(f):
    """Steps at which the fixture holds E(K), norms and the |k| <= 8 modes."""
    return sorted(int(k[len("low8_"):]) for k in f.files if k.startswith("low8_") and k[len("low8_"):].isdigit())


def test_les_256_fp64_500_steps_vs_oracle_fixture():
    """The whole 500-step horizon of the 256^3 LES against the FP64 oracle's
    run of the same seeded field (tests/golden/make_les256_fixture.py wrote
    the fixture on the host): time, dt, E, Omega and skewness every step,
    E(K), the component norms and the |k| <= 8 modes at checkpoints, all
    |k| <= 16 modes after step 500, each within relative 1e-12."""
    import spns_amd
    if not os.path.exists(FIXTURE):
        pytest.skip("oracle fixture not written")
    f = np.load(FIXTURE, allow_pickle=False)
    nu, cs, cat, delta, k0, seed, steps, N = f["params"]
    steps, N, seed = int(steps), int(N), int(seed)
    g = spns_amd.Solver(Nx=N, nu=nu, C_Smag=cs, C_At=cat, delta=delta)
    g.init_random(seed=seed, k0=k0)
    g.finish_init(False)
    P = float((3 * N // 2) ** 3)

    def rel(a, b):
        return abs(a - b) / abs(b)

    sc = g.scalars()
    worst = {"time": 0.0, "dt": rel(sc["dt"], f["dt"][0]), "ek": rel(sc["ek"], f["ek"][0]), "om": rel(sc["om"], f["om"][0]),
             "skew": 0.0, "ek_bins": 0.0, "norm": 0.0, "low8": 0.0, "low16": 0.0}
    cps = checkpoints(f)
    for n in range(1, steps + 1):
        g.step(1)
        sc = g.scalars()
        for k in ("time", "dt", "ek", "om"):
            worst[k] = max(worst[k], rel(sc[k], f[k][n]))
        sk = (f["s3"][n - 1] / P) / (f["s2"][n - 1] / P) ** 1.5
        worst["skew"] = max(worst["skew"], abs(sc["skewness"] - sk) / max(1.0, abs(sk)))
        if n in cps:
            st = g.get_state()
            ek = g.spectrum()
            ref = f["ek_bins_%d" % n]
            worst["ek_bins"] = max(worst["ek_bins"], float(np.linalg.norm(ek - ref) / np.linalg.norm(ref)))
            nrm = np.sum(np.abs(st) ** 2, axis=1)
            worst["norm"] = max(worst["norm"], float(np.max(np.abs(nrm - f["norm_%d" % n]) / f["norm_%d" % n])))
            lo = st[:, f["low8_idx"]]
            worst["low8"] = max(worst["low8"], float(np.linalg.norm(lo - f["low8_%d" % n]) / np.linalg.norm(f["low8_%d" % n])))
            if n == steps:
                lo = st[:, f["low_idx"]]
                worst["low16"] = float(np.linalg.norm(lo - f["low_%d" % n]) / np.linalg.norm(f["low_%d" % n]))
            print("les256 step %d: %s" % (n, {k: "%.1e" % v for k, v in worst.items()}), flush=True)
    for k, v in worst.items():
        assert v <= 1e-12, (k, v)
    g.close()
