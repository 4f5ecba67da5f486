"""FP32-split form (FP32 transforms and products, FP64 state, history,
reductions, forcing normalisation and time step) against the FP64 oracle.

Stated tolerances (the target is E(k) within 1e-5 of FP64):
E(k) rel L2 <= 1e-5, epsilon rel <= 1e-5, skewness abs <= 1e-4,
time and dt rel <= 1e-6, state rel L2 <= 1e-4."""
import numpy as np
import pytest

from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("case", [
    dict(Nx=64, nu=2e-3, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2),
    dict(Nx=32, nu=1e-4, C_Smag=0.17, C_At=0.5),
    dict(Nx=64, nu=1e-3, C_At=0.5, omega=(0.0, 0.0, 5.0)),
    # z transforms of the 512^3 and 1024^3 runs (768 and 1536 points, 257 and
    # 513 retained k3) on a small x-y cross-section
    dict(Nx=16, Ny=16, Nz=512, nu=2e-3, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2),
    dict(Nx=16, Ny=16, Nz=1024, nu=2e-3, C_At=0.5),
    dict(Nx=16, Mz=768, nu=2e-3, C_At=0.5, C_Smag=0.17),
])
def test_fp32_split_errors(case):
    import spns_amd
    g = spns_amd.Solver(precision=spns_amd.PREC_FP32_SPLIT, **case)
    o = Oracle(**case)
    g.init_random(seed=11, k0=4.0)
    o.set_state(g.get_state())
    g.finish_init(False)
    o.finish_init(False)
    g.step(20)
    o.step(20)
    sg, so = g.scalars(), o.scalars()
    ek_err = rel(g.spectrum(), o.spectrum())
    eps_err = abs(sg["nu"] * sg["om"] - so["nu"] * o.enstrophy()) / (so["nu"] * o.enstrophy())
    P = np.prod(o.shape)
    sk_o = (so["skew_s3"] / P) / (so["skew_s2"] / P) ** 1.5
    sk_err = abs(sg["skewness"] - sk_o)
    st_err = rel(g.get_state(), o.get_state())
    print("fp32-split errors: E(k) %.2e eps %.2e skew %.2e state %.2e" % (ek_err, eps_err, sk_err, st_err))
    assert ek_err <= 1e-5
    assert eps_err <= 1e-5
    assert sk_err <= 1e-4
    assert rel(sg["time"], so["time"]) <= 1e-6 and rel(sg["dt"], so["dt"]) <= 1e-6
    assert st_err <= 1e-4
