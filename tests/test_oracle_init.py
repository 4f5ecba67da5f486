"""The oracle's seeded synthetic initial field (the input of every measured
configuration; no reference counterpart).

What pins it, independent of the oracle's own code:
  * the hash: SplitMix64's published output sequence from state 0
    (Steele, Lea and Flood 2014; Vigna's splitmix64.c), and the uniform
    draws' range and first two moments;
  * the amplitudes: every mode's |u_hat|^2 is C * |k|^2 exp(-2 |k|^2/k0^2)
    with ONE constant C (the Box-Muller draws only set the direction), and
    C follows from E_tot = 1.5 u_rms^2 in closed form;
  * solenoidal per mode (k . u_hat = 0), Hermitian k3 = 0 plane (exactly),
    k = 0 mode zero, modes masked by the length factors zero;
  * the shell spectrum follows K^4 exp(-2 (K/k0)^2) up to shell counting;
  * isotropy and phases: each component holds about a third of the energy,
    the phases average out.
State layout: [component][i2][i1][i3] (the reference's Fourier-dump order).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, lib

GOLD = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


def wavenumbers(n):
    i = np.arange(n)
    return np.where(i <= n // 2, i, i - n)


def kgrid(o):
    nk2, nk1, nk3 = o.nk
    k2 = wavenumbers(nk2)[:, None, None]
    k1 = wavenumbers(nk1)[None, :, None]
    k3 = np.arange(nk3)[None, None, :]
    return np.broadcast_arrays(k1, k2, k3)


def test_splitmix_published_sequence():
    # splitmix64.c seeded with 0: the generator adds the golden gamma to its
    # state and mixes, so output n is the mix of n * gamma (+ gamma inside)
    L = lib()
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    got = [L.orc_splitmix((n * GOLD) & MASK) for n in range(4)]
    assert got == expect


def test_uniform_draws():
    L = lib()
    u = np.array([L.orc_rand_unif(7, k1, k2, k3, d) for k1 in range(-6, 7) for k2 in range(-6, 7)
                  for k3 in range(0, 7) for d in range(2)])
    assert np.all(u > 0) and np.all(u < 1)
    assert len(np.unique(u)) == len(u)
    n = len(u)
    assert abs(u.mean() - 0.5) < 4 * np.sqrt(1 / 12 / n)
    assert abs(u.var() - 1 / 12) < 0.01
    # keyed by the mode and the draw, not by call order
    assert L.orc_rand_unif(7, 3, -2, 5, 1) == L.orc_rand_unif(7, 3, -2, 5, 1)
    assert L.orc_rand_unif(7, 3, -2, 5, 1) != L.orc_rand_unif(8, 3, -2, 5, 1)
    assert L.orc_rand_unif(7, 3, -2, 5, 1) != L.orc_rand_unif(7, -3, 2, 5, 1)


@pytest.mark.parametrize("dims,k0,urms", [((16, 16, 16), 4.0, 1.0), ((12, 16, 20), 3.0, 0.7), ((32, 32, 32), 6.0, 1.0)])
def test_field_properties(dims, k0, urms):
    Nx, Ny, Nz = dims
    o = Oracle(Nx=Nx, Ny=Ny, Nz=Nz, nu=1e-2)
    o.init_random(seed=11, k0=k0, urms=urms)
    u = o.get_state().reshape(3, *o.nk)
    k1, k2, k3 = kgrid(o)
    r2 = (k1 * k1 + k2 * k2 + k3 * k3).astype(np.float64)
    w = np.where((k3 > 0) & ~((o.shape[2] % 2 == 0) & (k3 == o.shape[2] // 2)), 2.0, 1.0)
    e2 = np.sum(np.abs(u) ** 2, axis=0)
    # energy: E_tot = sum w |u|^2 / 2 = 1.5 urms^2
    E = 0.5 * np.sum(w * e2)
    assert abs(E - 1.5 * urms ** 2) <= 1e-14 * 1.5
    assert abs(o.energy() - E) <= 1e-13 * E
    # k = 0 zero; every other retained mode has |u|^2 = C k^2 exp(-2 k^2/k0^2)
    assert e2[0, 0, 0] == 0.0
    nz = r2 > 0
    shape = r2 * np.exp(-2.0 * r2 / k0 ** 2)
    C_closed = 1.5 * urms ** 2 / (0.5 * np.sum((w * shape)[nz]))
    ratio = e2[nz] / shape[nz]
    assert np.max(np.abs(ratio / C_closed - 1.0)) <= 1e-12
    # solenoidal per mode
    div = k1 * u[0] + k2 * u[1] + k3 * u[2]
    assert np.max(np.abs(div)) <= 1e-14 * np.max(np.sqrt(r2 * e2))
    # Hermitian k3 = 0 plane, exactly
    p = u[:, :, :, 0]
    j2 = (o.nk[0] - np.arange(o.nk[0])) % o.nk[0]
    j1 = (o.nk[1] - np.arange(o.nk[1])) % o.nk[1]
    assert np.array_equal(p, np.conj(p[:, j2][:, :, j1]))


def test_spectrum_shape_and_isotropy():
    o = Oracle(Nx=64, nu=1e-2)
    k0 = 6.0
    o.init_random(seed=3, k0=k0)
    ek = o.spectrum()
    K = np.arange(len(ek), dtype=np.float64)
    model = K ** 4 * np.exp(-2 * (K / k0) ** 2)
    sel = (K >= 3) & (K <= 20)
    c = np.sum(ek[sel]) / np.sum(model[sel])
    # shells hold about 4 pi K^2 modes (lattice counting noise, and the
    # exponential varies across a shell's width)
    assert np.max(np.abs(ek[sel] / (c * model[sel]) - 1.0)) < 0.25
    assert np.argmax(ek) in (int(k0) - 1, int(k0), int(k0) + 1)  # peak of K^4 exp(-2 K^2/k0^2) at K = k0
    # exactly: the shell sums of the closed-form per-mode amplitudes
    u = o.get_state().reshape(3, *o.nk)
    k1, k2, k3 = kgrid(o)
    r2 = (k1 * k1 + k2 * k2 + k3 * k3).astype(np.float64)
    w = np.where(k3 > 0, 2.0, 1.0)
    per = r2 * np.exp(-2.0 * r2 / k0 ** 2)
    C = 1.5 / (0.5 * np.sum(w * per))
    shells = np.rint(np.sqrt(r2)).astype(int)
    ek_np = np.bincount(shells.ravel(), weights=(0.5 * w * C * per).ravel(), minlength=len(ek))
    assert np.max(np.abs(ek - ek_np[:len(ek)])) <= 1e-13 * np.max(ek)
    ec = np.array([0.5 * np.sum(w * np.abs(u[c]) ** 2) for c in range(3)])
    assert np.max(np.abs(ec / ec.sum() - 1 / 3)) < 0.02
    # random phases: the mean unit phasor over many modes is near zero
    ph = u[0][np.abs(u[0]) > 0]
    ph = ph / np.abs(ph)
    assert abs(ph.mean()) < 5 / np.sqrt(ph.size)


def test_masked_modes_zero():
    # 1:1:2 box with length factors (2, 2, 1): only k_i divisible by L_i
    o = Oracle(Nx=32, Ny=32, Nz=16, Mx=48, My=48, Mz=24, Lx=2, Ly=2, Lz=1, isNotCubic=True, nu=1e-3)
    o.init_random(seed=5, k0=4.0)
    u = o.get_state().reshape(3, *o.nk)
    k1, k2, k3 = kgrid(o)
    masked = (k1 % 2 != 0) | (k2 % 2 != 0)
    assert np.all(u[:, masked] == 0)
    assert np.count_nonzero(u[0][~masked]) > 0
    assert abs(o.energy() - 1.5) <= 1e-14 * 1.5


def test_seed_changes_phases_not_amplitudes():
    a = Oracle(Nx=16, nu=1e-2)
    b = Oracle(Nx=16, nu=1e-2)
    a.init_random(seed=1, k0=4.0)
    b.init_random(seed=2, k0=4.0)
    ua, ub = a.get_state(), b.get_state()
    ea = np.sum(np.abs(ua.reshape(3, -1)) ** 2, axis=0)
    eb = np.sum(np.abs(ub.reshape(3, -1)) ** 2, axis=0)
    assert np.allclose(ea, eb, rtol=1e-12, atol=0)
    assert np.linalg.norm(ua - ub) > 0.5 * np.linalg.norm(ua)
