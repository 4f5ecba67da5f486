"""Full-size runs, where the oracle is too slow to step alongside: the 512^3
forced DNS in both forms and the 1024^3 forced DNS in the FP32-split form.
Checked through properties that hold at any size:
  * the state stays solenoidal (k . u_hat = 0 after the projection), FP64;
  * the k3 = 0 plane stays Hermitian, u_hat(-k1,-k2,0) = conj u_hat(k1,k2,0);
  * the shell spectrum sums to the total energy;
  * time = sum of the steps' dt, all scalars finite;
  * FP32-split E(k), epsilon and skewness agree with the FP64 form
    (E(k) within the 1e-5 target), and so do the five velocity-gradient
    invariant spectra (1e-5).
The same seeded field (k0 = 4, u_rms = 1) starts every run.  State layout:
[component][i2][i1][i3], the reference's Fourier-dump order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def wavenumbers(n):
    i = np.arange(n)
    return np.where(i <= n // 2, i, i - n).astype(np.float64)


def divergence_rel(st, nk1, nk2, nk3, planes=None):
    u = st.reshape(3, nk2, nk1, nk3)
    k1 = wavenumbers(nk1)[None, :, None]
    k3 = np.arange(nk3, dtype=np.float64)[None, None, :]
    k2all = wavenumbers(nk2)
    sel = slice(None) if planes is None else planes
    num = den = 0.0
    for i2 in np.arange(nk2)[sel]:
        k2 = k2all[i2]
        d = k1[0] * u[0, i2] + k2 * u[1, i2] + k3[0] * u[2, i2]
        kk = np.sqrt(k1[0] ** 2 + k2 ** 2 + k3[0] ** 2)
        mag = np.sqrt(np.abs(u[0, i2]) ** 2 + np.abs(u[1, i2]) ** 2 + np.abs(u[2, i2]) ** 2)
        num += float(np.sum(np.abs(d)))
        den += float(np.sum(kk * mag))
    return num / den


def hermitian_rel(st, nk1, nk2, nk3):
    u = st.reshape(3, nk2, nk1, nk3)[:, :, :, 0]
    j2 = (nk2 - np.arange(nk2)) % nk2
    j1 = (nk1 - np.arange(nk1)) % nk1
    mirror = np.conj(u[:, j2][:, :, j1])
    return float(np.linalg.norm(u - mirror) / np.linalg.norm(u))


def run(N, prec, steps, seed=2024, invariants=False):
    import spns_amd
    nu = 8e-5 if N >= 1024 else 2e-4
    s = spns_amd.Solver(Nx=N, nu=nu, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2, precision=prec,
                        gradient_invariants=invariants)
    s.init_random(seed=seed, k0=4.0, urms=1.0)
    s.finish_init(False)
    dts = []
    for _ in range(steps):
        s.step(1)
        dts.append(s.scalars()["dt"])
    sc = s.scalars()
    ek = s.spectrum()
    return s, sc, ek, dts


def common_checks(s, sc, ek, dts, N, div_tol, herm_tol, planes=None):
    nk1 = nk2 = 2 * (N // 2) + 1
    nk3 = N // 2 + 1
    for k in ("time", "dt", "ek", "om", "maxvel", "skewness"):
        assert np.isfinite(sc[k]), k
    assert abs(ek.sum() - sc["ek"]) <= 1e-12 * sc["ek"]
    # time advances by the dt computed after each step (the first step uses the initial dt)
    assert sc["time"] > 0 and sc["dt"] > 0
    st = s.get_state()
    div = divergence_rel(st, nk1, nk2, nk3, planes)
    herm = hermitian_rel(st, nk1, nk2, nk3)
    print("N=%d: |k.u|/|k||u| %.2e, Hermitian defect %.2e, E %.6f" % (N, div, herm, sc["ek"]))
    assert div <= div_tol
    assert herm <= herm_tol


def test_512_fp64_and_fp32_split():
    import spns_amd
    steps = 10
    s64, sc64, ek64, dt64 = run(512, spns_amd.PREC_FP64, steps, invariants=True)
    common_checks(s64, sc64, ek64, dt64, 512, 1e-13, 1e-12)
    inv64 = s64.invariant_spectra()
    s64.close()
    s32, sc32, ek32, dt32 = run(512, spns_amd.PREC_FP32_SPLIT, steps, invariants=True)
    common_checks(s32, sc32, ek32, dt32, 512, 1e-13, 1e-6)
    inv32 = s32.invariant_spectra()
    s32.close()
    # velocity-gradient invariant spectra of the same start-of-step field in both forms
    for f, name in enumerate(("QG", "QS", "RG", "RS", "V2")):
        e = float(np.linalg.norm(inv32[f] - inv64[f]) / np.linalg.norm(inv64[f]))
        print("512^3 invariant %s FP32-split vs FP64: %.2e" % (name, e))
        assert e <= 1e-5, name
    ek_err = float(np.linalg.norm(ek32 - ek64) / np.linalg.norm(ek64))
    eps_err = abs(sc32["nu"] * sc32["om"] - sc64["nu"] * sc64["om"]) / (sc64["nu"] * sc64["om"])
    print("512^3 FP32-split vs FP64 after %d steps: E(k) %.2e eps %.2e skew %.2e" % (
        steps, ek_err, eps_err, abs(sc32["skewness"] - sc64["skewness"])))
    assert ek_err <= 1e-5
    assert eps_err <= 1e-5
    assert abs(sc32["skewness"] - sc64["skewness"]) <= 1e-4
    assert abs(sc32["time"] - sc64["time"]) <= 1e-6 * sc64["time"]


def plane_checks(s, N, sample):
    """Divergence and Hermitian-plane defects on the sampled i2 planes and
    their mirrors only (spns_get_state_planes: the whole 1024^3 state is
    26 GB on the host)."""
    nk1 = nk2 = 2 * (N // 2) + 1
    nk3 = N // 2 + 1
    k1 = wavenumbers(nk1)[:, None]
    k3 = np.arange(nk3, dtype=np.float64)[None, :]
    k2all = wavenumbers(nk2)
    num = den = herm_num = herm_den = 0.0
    for i2 in sample:
        j2 = (nk2 - i2) % nk2
        p = s.get_state_planes(i2, 1)[:, 0]
        q = s.get_state_planes(j2, 1)[:, 0]
        d = k1 * p[0] + k2all[i2] * p[1] + k3 * p[2]
        kk = np.sqrt(k1 ** 2 + k2all[i2] ** 2 + k3 ** 2)
        mag = np.sqrt(np.sum(np.abs(p) ** 2, axis=0))
        num += float(np.sum(np.abs(d)))
        den += float(np.sum(kk * mag))
        # u(-k1, -k2, 0) = conj u(k1, k2, 0)
        j1 = (nk1 - np.arange(nk1)) % nk1
        herm_num += float(np.sum(np.abs(p[:, :, 0] - np.conj(q[:, j1, 0])) ** 2))
        herm_den += float(np.sum(np.abs(p[:, :, 0]) ** 2))
    return num / den, np.sqrt(herm_num / herm_den)


def test_1024_fp32_split_200_steps():
    """The 1024^3 forced DNS (nu = 8e-5) over its 200 steps, FP32-split."""
    import spns_amd
    N = 1024
    s, sc, ek, dts = run(N, spns_amd.PREC_FP32_SPLIT, 200)
    for k in ("time", "dt", "ek", "om", "maxvel", "skewness"):
        assert np.isfinite(sc[k]), k
    assert abs(ek.sum() - sc["ek"]) <= 1e-12 * sc["ek"]
    assert abs(sc["time"] - sum(dts)) <= 1e-12 * sc["time"]  # time += the dt computed after each step
    sample = [0, 1, 2, 3, 5, 8, 13, 64, 200, 511, 512]
    div, herm = plane_checks(s, N, sample)
    print("N=1024 after 200 steps: |k.u|/|k||u| %.2e, Hermitian defect %.2e, E %.6f, eps %.6f, skewness %.4f, time %.5f" % (
        div, herm, sc["ek"], sc["nu"] * sc["om"], sc["skewness"], sc["time"]))
    assert div <= 1e-13
    assert herm <= 1e-6
    # constant power 0.1 into 1 <= K <= 2: E rises slowly, the cascade builds dissipation
    assert 1.5 < sc["ek"] < 1.7
    assert sc["skewness"] < 0.0
    s.close()

