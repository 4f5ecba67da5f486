"""The 256^3 Smagorinsky LES at full size, FP64 plain form against the FP64
oracle: k0 = 6 seeded field (u_rms = 1), nu = 1e-4, C_s = 0.17,
Delta = 2 pi / 256, CFL step with C_At = 0.5.  A few steps (the oracle takes
seconds per step at this size on the host); state, R history, E(k), time,
dt, energy and enstrophy within relative L2 1e-12 after each chunk.  The
skewness partial sums agree as well."""
import math

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
