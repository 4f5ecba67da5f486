"""Pins the oracle's parts by mathematics independent of the reference logs:
FFTs vs numpy and brute-force DFT, Parseval (src/HIT_LES.cpp:1216-1252),
the AB2 viscous-decay recurrence including the R0=0 first step
(src/HIT_LES.cpp:322-338), the CFL closed form (src/HIT_LES.cpp:389-403),
projection (src/HIT_LES.cpp:346-359) and the constant-power forcing law
(src/HIT_LES.cpp:243-278)."""
import numpy as np
import pytest

from oracle import oracle as O


def compact_index(nk, k1, k2, k3):
    nk2, nk1, nk3 = nk
    i2 = k2 if k2 >= 0 else nk2 + k2
    i1 = k1 if k1 >= 0 else nk1 + k1
    return (i2 * nk1 + i1) * nk3 + k3


def random_state(o, seed=0, scale=1.0):
    """Random retained field, Hermitian on k3=0, solenoidal, zero mean."""
    rng = np.random.default_rng(seed)
    nk2, nk1, nk3 = o.nk
    st = (rng.standard_normal((3, o.nmodes)) + 1j * rng.standard_normal((3, o.nmodes))) * scale
    # impose Hermitian symmetry on k3=0 through a physical round trip in numpy
    o.set_state(st)
    o.antitransform()
    u = o.velocity_physical()
    Mx, My, Mz = o.shape
    spec = np.fft.rfftn(u, axes=(1, 2, 3)) / (Mx * My * Mz)  # [c][x][y][z]
    k1 = np.fft.fftfreq(Mx, 1.0 / Mx).astype(int)
    k2 = np.fft.fftfreq(My, 1.0 / My).astype(int)
    out = np.zeros((3, o.nmodes), dtype=complex)
    Nx2, Ny2, Nz2 = nk1 // 2, nk2 // 2, nk3 - 1
    for kk2 in range(-Ny2, Ny2 + 1):
        for kk1 in range(-Nx2, Nx2 + 1):
            for kk3 in range(0, Nz2 + 1):
                m = compact_index(o.nk, kk1, kk2, kk3)
                out[:, m] = spec[:, kk1 % Mx, kk2 % My, kk3]
    # solenoidal projection and zero mean
    for kk2 in range(-Ny2, Ny2 + 1):
        for kk1 in range(-Nx2, Nx2 + 1):
            for kk3 in range(0, Nz2 + 1):
                m = compact_index(o.nk, kk1, kk2, kk3)
                k = np.array([kk1, kk2, kk3], dtype=float)
                r2 = k @ k
                if r2 == 0:
                    out[:, m] = 0
                else:
                    out[:, m] -= k * (k @ out[:, m]) / r2
    del k1, k2
    o.set_state(out)
    return out * np.sqrt(1.5 / o.energy())  # u_rms = 1


@pytest.mark.parametrize("n", [8, 12, 16, 24, 36, 48, 72, 96, 144, 192, 288, 384, 768])
def test_fft1d_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(O.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-13 * np.sqrt(n) * np.abs(x).max() * 4)
    np.testing.assert_allclose(O.fft1d(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-13 * np.sqrt(n) * np.abs(x).max() * 4)


def test_fft1d_brute_force():
    n = 12
    x = np.arange(n) * 0.37 + 1j * np.cos(np.arange(n))
    W = np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
    np.testing.assert_allclose(O.fft1d(x, -1), W @ x, atol=1e-13)


@pytest.mark.parametrize("shape", [(6, 8, 12), (12, 18, 24), (48, 48, 48), (24, 12, 18)])
def test_fft3d_r2c_c2r(shape):
    Mx, My, Mz = shape
    rng = np.random.default_rng(1)
    r = rng.standard_normal(shape)
    rp = np.zeros((Mx, My, 2 * (Mz // 2 + 1)))
    rp[:, :, :Mz] = r
    c = O.fft_r2c(rp, Mx, My, Mz)
    ref = np.fft.rfftn(r).transpose(1, 0, 2)  # reference layout [y][x][k3]
    np.testing.assert_allclose(c, ref, atol=1e-12 * np.sqrt(r.size))
    back = O.fft_c2r(c, Mx, My, Mz)[:, :, :Mz]
    np.testing.assert_allclose(back, np.fft.irfftn(np.fft.rfftn(r), s=shape, axes=(0, 1, 2)) * r.size, atol=1e-11)
    np.testing.assert_allclose(back / r.size, r, atol=1e-14)


def test_fft3d_brute_force_small():
    Mx, My, Mz = 4, 6, 6
    rng = np.random.default_rng(3)
    r = rng.standard_normal((Mx, My, Mz))
    rp = np.zeros((Mx, My, 2 * (Mz // 2 + 1)))
    rp[:, :, :Mz] = r
    c = O.fft_r2c(rp, Mx, My, Mz)
    x = np.arange(Mx)[:, None, None]
    y = np.arange(My)[None, :, None]
    z = np.arange(Mz)[None, None, :]
    for b in range(My):
        for a in range(Mx):
            for k3 in range(Mz // 2 + 1):
                e = np.exp(-2j * np.pi * (a * x / Mx + b * y / My + k3 * z / Mz))
                assert abs(c[b, a, k3] - np.sum(r * e)) < 1e-12


def test_c2r_ignores_non_hermitian_imag_of_k3_zero_dc():
    """A c2r of a 3-D field does the real dimension last, so Im of the
    (0,0,0) coefficient never reaches the output."""
    Mx = My = Mz = 12
    c = np.zeros((My, Mx, Mz // 2 + 1), dtype=complex)
    c[0, 0, 0] = 1.0 + 5.0j
    out = O.fft_c2r(c, Mx, My, Mz)[:, :, :Mz]
    np.testing.assert_allclose(out, np.ones((Mx, My, Mz)), atol=1e-15)


def test_parseval_random_state():
    o = O.Oracle(16, nu=0.01)
    o.set_state(random_state(o, 5))
    ekf, ekr, eks = o.parseval()
    assert abs(ekf - ekr) < 1e-13 and abs(eks - ekf) < 1e-13


def _shear_state(o, amp, k2=1):
    st = np.zeros((3, o.nmodes), dtype=complex)
    st[0, compact_index(o.nk, 0, k2, 0)] = amp
    st[0, compact_index(o.nk, 0, -k2, 0)] = np.conj(amp)
    return st


def test_ab2_viscous_decay_recurrence():
    """u_x = 2A cos(k y) has u.grad u = 0: the step is pure viscous AB2
    decay, u1 = u0 - 1.5 dt nu k^2 u0 (R0=0), then
    u_{n+1} = u_n + dt(-1.5 nu k^2 u_n + 0.5 nu k^2 u_{n-1})."""
    nu, dt, k2 = 0.05, 1e-2, 2
    o = O.Oracle(16, nu=nu, fixed_dt=dt)
    o.set_state(_shear_state(o, 0.3, k2))
    o.finish_init(False)
    a = [0.3]
    lam = nu * k2 * k2
    a.append(a[0] + dt * (1.5 * (-lam * a[0])))
    for n in range(1, 6):
        a.append(a[n] + dt * (1.5 * (-lam * a[n]) - 0.5 * (-lam * a[n - 1])))
    for n in range(1, 7):
        o.step()
        got = o.get_state()[0, compact_index(o.nk, 0, k2, 0)]
        assert abs(got - a[n]) < 1e-15, (n, got, a[n])


def test_cfl_closed_form():
    nu, C_At, A = 0.01, 0.5, 0.7
    o = O.Oracle(16, nu=nu, C_At=C_At)
    o.set_state(_shear_state(o, A))
    o.finish_init(False)
    Ax = 2 * np.pi / 24
    maxvel = 2 * A / Ax
    dt = C_At * min(0.2 * Ax * Ax / nu, 0.35 / maxvel)
    s = o.scalars()
    assert abs(s["maxvel"] - maxvel) < 1e-12 * maxvel
    assert abs(s["dt"] - dt) < 1e-15 * dt


def test_projection_divergence_free_after_step():
    o = O.Oracle(16, nu=0.01, fixed_dt=1e-3)
    st = random_state(o, 7)
    rng = np.random.default_rng(8)
    st = st + 0.1 * (rng.standard_normal(st.shape) + 1j * rng.standard_normal(st.shape))  # not solenoidal
    o.set_state(st)
    o.finish_init(False)
    o.step()
    s = o.get_state()
    nk2, nk1, nk3 = o.nk
    worst = 0.0
    for k2 in range(-(nk2 // 2), nk2 // 2 + 1):
        for k1 in range(-(nk1 // 2), nk1 // 2 + 1):
            for k3 in range(nk3):
                m = compact_index(o.nk, k1, k2, k3)
                div = k1 * s[0, m] + k2 * s[1, m] + k3 * s[2, m]
                worst = max(worst, abs(div) / (1 + np.abs(s[:, m]).sum() * np.sqrt(k1 * k1 + k2 * k2 + k3 * k3)))
    assert worst < 1e-15


def test_constant_power_first_step():
    """nu=0, one shear mode at K=1 (no nonlinearity): R = alpha u with
    alpha = eps/(2E), so E1 = E0 (1 + 1.5 dt alpha)^2 on the R0=0 step."""
    eps, dt = 0.1, 1e-2
    o = O.Oracle(16, nu=1e-300, fixed_dt=dt, force_eps=eps, force_kmin=1, force_kmax=2)
    o.set_state(_shear_state(o, 0.25, 1))
    o.finish_init(False)
    e0 = o.energy()
    alpha = eps / (2 * e0)
    o.step()
    assert abs(o.energy() - e0 * (1 + 1.5 * dt * alpha) ** 2) < 1e-15


def test_spectrum_sums_to_energy():
    o = O.Oracle(16, nu=0.01)
    o.set_state(random_state(o, 11))
    ek = o.spectrum()
    assert abs(ek.sum() - o.energy()) < 1e-14 * o.energy()
    assert ek[0] == 0.0


def test_les_filter_width_enters_as_cs_times_delta():
    """The explicit LES filter width (an option; the reference always uses
    2 pi / cbrt(Mx My Mz), src/HIT_LES.h:53) enters only through
    nu_t = (C_s Delta)^2 |S| (src/HIT_LES.cpp:130-142): the reference's
    Delta given explicitly reproduces the default run bit for bit, the same
    product C_s Delta reproduces it to rounding, and a larger product
    dissipates more."""
    d_ref = 2 * np.pi / 48  # 32^3 retained modes, FFT grid 48^3
    kw = dict(nu=1e-4, C_At=0.5)
    runs = {"default": O.Oracle(32, C_Smag=0.17, **kw),
            "explicit": O.Oracle(32, C_Smag=0.17, delta=d_ref, **kw),
            "same_product": O.Oracle(32, C_Smag=0.17 * 1.7, delta=d_ref / 1.7, **kw),
            "larger": O.Oracle(32, C_Smag=0.17, delta=1.5 * d_ref, **kw)}
    for o in runs.values():
        random_state(o, 3)
        o.finish_init(False)
        o.step(5)
    a = runs["default"].get_state()
    assert np.array_equal(runs["explicit"].get_state(), a)
    c = runs["same_product"].get_state()
    assert np.linalg.norm(c - a) <= 1e-13 * np.linalg.norm(a)
    assert runs["larger"].energy() < runs["default"].energy()
