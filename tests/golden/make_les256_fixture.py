"""Writes tests/golden/les256_k0_6_seed6.npz: the FP64 oracle's run of the
256^3 Smagorinsky LES over its full horizon, for the GPU parity test.

Setup (the LES configuration): N = 256 (FFT grid 384^3), seeded field with
E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 6, u_rms = 1, seed 6; nu = 1e-4,
C_s = 0.17, Delta = 2 pi / 256, CFL step with C_At = 0.5; 500 steps.

Stored (plain arrays; numpy.load with allow_pickle=False reads them):
  params            nu, C_s, C_At, Delta, k0, seed, steps, N
  time, dt, ek, om  per step 0..500 (row 0 after the initial CFL step)
  s2, s3            per step 1..500: sums of (du/dx)^2 and (du/dx)^3 over the
                    384^3 grid, of the field at the start of the step
  ek_bins_<n>       E(K), K = 0..last_rad_max, after step n
  low_idx           compact indices [i2][i1][i3] of the retained modes with
                    |k| <= 16 (the energy-containing range)
  low_<n>           the state at those modes after step n, [3][len(low_idx)]
  low8_idx, low8_<n> the same for |k| <= 8 at the intermediate checkpoints
  norm_<n>          sum over the compact box of |u_hat_c|^2, per component
The oracle takes seconds per step at this size; the whole run is about an
hour on 8 host cores.  Run from the repository root:
    python tests/golden/make_les256_fixture.py [steps]
"""
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402

N = 256
SEED = 6
K0 = 6.0
KW = dict(Nx=N, nu=1e-4, C_Smag=0.17, C_At=0.5, delta=2.0 * math.pi / 256)
CHECKPOINTS = (1, 10, 50, 100, 200, 300, 400, 500)


def low_index(nk, kmax):
    nk2, nk1, nk3 = nk
    w = lambda n: np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n)
    k2 = w(nk2)[:, None, None]
    k1 = w(nk1)[None, :, None]
    k3 = np.arange(nk3)[None, None, :]
    r2 = k1 * k1 + k2 * k2 + k3 * k3
    return np.flatnonzero((r2 <= kmax * kmax).ravel())


def main(steps=500):
    out = {"params": np.array([KW["nu"], KW["C_Smag"], KW["C_At"], KW["delta"], K0, SEED, steps, N], dtype=np.float64)}
    o = Oracle(**KW)
    o.init_random(seed=SEED, k0=K0)
    o.finish_init(False)
    idx16 = low_index(o.nk, 16)
    idx8 = low_index(o.nk, 8)
    out["low_idx"] = idx16.astype(np.int64)
    out["low8_idx"] = idx8.astype(np.int64)
    rows = {k: [] for k in ("time", "dt", "ek", "om", "s2", "s3")}

    def record():
        sc = o.scalars()
        rows["time"].append(sc["time"])
        rows["dt"].append(sc["dt"])
        rows["ek"].append(o.energy())
        rows["om"].append(o.enstrophy())

    record()
    t0 = time.time()
    for n in range(1, steps + 1):
        o.step(1)
        record()
        sc = o.scalars()
        rows["s2"].append(sc["skew_s2"])
        rows["s3"].append(sc["skew_s3"])
        if n in CHECKPOINTS or n == steps:
            st = o.get_state()
            out["low8_%d" % n] = st[:, idx8].copy()
            out["norm_%d" % n] = np.sum(np.abs(st) ** 2, axis=1)
            out["ek_bins_%d" % n] = o.spectrum()
            if n == steps:
                out["low_%d" % n] = st[:, idx16].copy()
        print("step %d/%d  t=%.6f dt=%.3e E=%.9f  %.1f s/step" % (n, steps, rows["time"][-1], rows["dt"][-1],
                                                               rows["ek"][-1], (time.time() - t0) / n), flush=True)
    for k, v in rows.items():
        out[k] = np.array(v, dtype=np.float64)
    path = os.path.join(HERE, "les256_k0_6_seed6.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(*(int(a) for a in sys.argv[1:]))
