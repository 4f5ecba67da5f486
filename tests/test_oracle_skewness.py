"""Velocity-gradient skewness sums of the oracle, pinned independently of its
code.  The reference computes no skewness; the quantity is
S = <(du/dx)^3> / <(du/dx)^2>^(3/2) over the M grid, with du/dx the physical
field of i k1 u_x-hat at the start of the step (the gradient field gradu[0]
that the reference's convective term forms, src/HIT_LES.cpp:114-116).  The
oracle reports the two sums sum (du/dx)^2 and sum (du/dx)^3 of the last step.

  * numpy restatement: random solenoidal states (cubic and 12x16x20), du/dx
    by numpy's irfftn of i k1 u_x-hat, the sums compared at 1e-12 relative,
    after one step and after three (the sums follow the start of the last
    step, not the end);
  * closed forms: a pure shear u_x(y) has du/dx = 0; the 2-D cellular flow
    with stream function sin x sin y + b sin 2x sin 2y has
    <(du/dx)^2> = 1/4 + 4 b^2 and <(du/dx)^3> = 3b/4 exactly on the grid.
    A wrong component (dv/dx, du/dy) or a wrong moment fails one of them.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle


def wrap(i, n):
    return np.where(i <= n // 2, i, i - n)


def compact_to_full(state, N3, M3):
    """compact [3][nk2][nk1][nk3] -> full half spectra [3][My][Mx][Mz//2+1]."""
    Nx, Ny, Nz = N3
    Mx, My, Mz = M3
    nk1, nk2, nk3 = 2 * (Nx // 2) + 1, 2 * (Ny // 2) + 1, Nz // 2 + 1
    u = state.reshape(3, nk2, nk1, nk3)
    full = np.zeros((3, My, Mx, Mz // 2 + 1), complex)
    k1 = wrap(np.arange(nk1), nk1)
    k2 = wrap(np.arange(nk2), nk2)
    full[:, (k2 % My)[:, None, None], (k1 % Mx)[None, :, None], np.arange(nk3)[None, None, :]] = u
    return full


def full_to_compact(full, N3, M3):
    Nx, Ny, Nz = N3
    Mx, My, Mz = M3
    nk1, nk2, nk3 = 2 * (Nx // 2) + 1, 2 * (Ny // 2) + 1, Nz // 2 + 1
    k1 = wrap(np.arange(nk1), nk1)
    k2 = wrap(np.arange(nk2), nk2)
    return full[:, (k2 % My)[:, None, None], (k1 % Mx)[None, :, None], np.arange(nk3)[None, None, :]].reshape(3, -1)


def physical_to_compact(uvw, N3, M3):
    """uvw [3][My][Mx][Mz] (physical, y outer) -> normalised compact state."""
    Mx, My, Mz = M3
    full = np.fft.rfftn(uvw, axes=(1, 2, 3)) / (Mx * My * Mz)
    return full_to_compact(full, N3, M3)


def dudx(state, N3, M3):
    """du/dx on the M grid ([y][x][z]) from i k1 u_x-hat, by numpy."""
    Mx, My, Mz = M3
    full = compact_to_full(state, N3, M3)[0]
    kx = np.fft.fftfreq(Mx, 1.0 / Mx)[None, :, None]
    return np.fft.irfftn(1j * kx * full, s=(My, Mx, Mz), axes=(0, 1, 2)) * (Mx * My * Mz)


def random_solenoidal(N3, M3, seed):
    """Random real field on the M grid, transformed, truncated to the
    retained box (so the k3 = 0 plane is Hermitian) and projected."""
    Mx, My, Mz = M3
    rng = np.random.default_rng(seed)
    st = physical_to_compact(rng.standard_normal((3, My, Mx, Mz)), N3, M3)
    Nx, Ny, Nz = N3
    nk1, nk2, nk3 = 2 * (Nx // 2) + 1, 2 * (Ny // 2) + 1, Nz // 2 + 1
    k2, k1, k3 = np.meshgrid(wrap(np.arange(nk2), nk2), wrap(np.arange(nk1), nk1), np.arange(nk3), indexing="ij")
    k = np.stack([k1.ravel(), k2.ravel(), k3.ravel()]).astype(float)
    kk = (k * k).sum(0)
    kk[kk == 0] = 1.0
    st = st - k * (k * st).sum(0) / kk
    return st * 3.0  # order-one gradients


CASES = [((16, 16, 16), dict(Nx=16)), ((12, 16, 20), dict(Nx=12, Ny=16, Nz=20))]


@pytest.mark.parametrize("N3,kw", CASES)
def test_skew_sums_against_numpy(N3, kw):
    M3 = tuple((3 * n) // 2 for n in N3)
    o = Oracle(nu=1e-2, C_At=0.5, **kw)
    # a random solenoidal field plus a cellular flow with a skewed du/dx
    o.set_state(random_solenoidal(N3, M3, seed=11) + 4.0 * _cellular_state(N3, M3, 0.3))
    o.finish_init(False)
    start = o.get_state()
    o.step(1)
    g = dudx(start, N3, M3)
    ref2, ref3 = np.sum(g ** 2), np.sum(g ** 3)
    sc = o.scalars()
    assert abs(sc["skew_s2"] - ref2) <= 1e-12 * abs(ref2)
    assert abs(sc["skew_s3"] - ref3) <= 1e-12 * np.sum(np.abs(g) ** 3)
    P = g.size
    assert abs((ref3 / P) / (ref2 / P) ** 1.5) > 0.02  # a non-trivial third moment
    # three steps in all: the sums belong to the start of the third step
    o.step(1)
    s_before_last = o.get_state()
    o.step(1)
    g = dudx(s_before_last, N3, M3)
    sc = o.scalars()
    assert abs(sc["skew_s2"] - np.sum(g ** 2)) <= 1e-12 * np.sum(g ** 2)
    assert abs(sc["skew_s3"] - np.sum(g ** 3)) <= 1e-12 * np.sum(np.abs(g) ** 3)
    g_end = dudx(o.get_state(), N3, M3)
    assert abs(sc["skew_s2"] - np.sum(g_end ** 2)) > 1e-9 * np.sum(g ** 2)


def _cellular_state(N3, M3, b):
    """psi = sin x sin y + b sin 2x sin 2y; u = d psi / dy, v = -d psi / dx, w = 0."""
    Mx, My, Mz = M3
    X = 2 * np.pi * np.arange(Mx)[None, :, None] / Mx
    Y = 2 * np.pi * np.arange(My)[:, None, None] / My
    u = np.sin(X) * np.cos(Y) + 2 * b * np.sin(2 * X) * np.cos(2 * Y)
    v = -np.cos(X) * np.sin(Y) - 2 * b * np.cos(2 * X) * np.sin(2 * Y)
    uvw = np.zeros((3, My, Mx, Mz))
    uvw[0] += u
    uvw[1] += v
    return physical_to_compact(uvw, N3, M3)


@pytest.mark.parametrize("b", [0.0, 0.3, -0.45])
def test_skew_sums_closed_form_cellular(b):
    N = 16
    M = 24
    st = _cellular_state((N, N, N), (M, M, M), b)
    o = Oracle(Nx=N, nu=1e-2, fixed_dt=1e-4)
    o.set_state(st)
    o.finish_init(False)
    o.step(1)
    sc = o.scalars()
    P = M ** 3
    assert abs(sc["skew_s2"] - P * (0.25 + 4 * b * b)) <= 1e-12 * P
    assert abs(sc["skew_s3"] - P * 0.75 * b) <= 1e-12 * P
    S = sc["skew_s3"] / P / (sc["skew_s2"] / P) ** 1.5
    assert abs(S - 0.75 * b / (0.25 + 4 * b * b) ** 1.5) <= 1e-12


def test_skew_sums_closed_form_shear():
    """u_x = sin y + 0.5 cos 3y: du/dx = 0, both sums vanish (a component
    mix-up with du/dy would not)."""
    N = 16
    M = 24
    y = 2 * np.pi * np.arange(M) / M
    Y = np.broadcast_to(y[:, None, None], (M, M, M))
    u = np.sin(Y) + 0.5 * np.cos(3 * Y)
    st = physical_to_compact(np.stack([u, np.zeros_like(u), np.zeros_like(u)]), (N, N, N), (M, M, M))
    o = Oracle(Nx=N, nu=1e-2, fixed_dt=1e-4)
    o.set_state(st)
    o.finish_init(False)
    o.step(1)
    sc = o.scalars()
    assert abs(sc["skew_s2"]) <= 1e-20 * M ** 3
    assert abs(sc["skew_s3"]) <= 1e-20 * M ** 3
