"""Velocity-gradient invariant spectra of the oracle (src/HIT_LES.cpp:1297-1319,
src/tensor_algebra.cpp:250-303), pinned independently of the oracle's code:
  * the per-point invariants against numpy (traces, determinants, matrix
    products) on random tensors, and closed forms for a pure rotation and a
    pure strain;
  * the five shell spectra against a numpy restatement (numpy FFTs of the
    gradient fields, numpy invariants, rfftn, the D mask, the conjugate
    weights and the integer shell rounding);
  * after steps, the spectra follow the state at the start of the last step
    (the field whose gradients the reference keeps after a step)."""
import numpy as np
import pytest

from oracle.oracle import Oracle, invariants_point


def np_invariants(G):
    """G [..., 3, 3], G[i][j] = d u_i / d x_j -> QG QS RG RS V2."""
    S = 0.5 * (G + np.swapaxes(G, -1, -2))
    O = G - S
    tr = lambda A: np.trace(A, axis1=-2, axis2=-1)
    q = lambda A: 0.5 * (tr(A) ** 2 - tr(A @ A))
    QG, QS = q(G), q(S)
    V2 = 4.0 * (tr(S @ S @ O @ O) - 2.0 * QS * (QG - QS))
    return np.stack([QG, QS, np.linalg.det(G), np.linalg.det(S), V2], axis=-1)


def test_point_invariants_random():
    rng = np.random.default_rng(3)
    for _ in range(50):
        G = rng.standard_normal((3, 3))
        assert np.allclose(invariants_point(G), np_invariants(G), rtol=1e-13, atol=1e-13)


def test_point_invariants_closed_forms():
    w = 1.7
    rot = np.array([[0.0, -w, 0.0], [w, 0.0, 0.0], [0.0, 0.0, 0.0]])
    assert np.allclose(invariants_point(rot), [w * w, 0.0, 0.0, 0.0, 0.0], atol=1e-14)
    a, b = 0.3, -1.1
    st = np.diag([a, b, -a - b])
    q = -0.5 * (a * a + b * b + (a + b) ** 2)
    r = -a * b * (a + b)
    assert np.allclose(invariants_point(st), [q, q, r, r, 0.0], atol=1e-14)
    # transposing G keeps S and flips Omega's sign: all five invariants are unchanged;
    # a trace changes QG and RG
    G = np.array([[0.1, 0.7, -0.2], [0.0, -0.4, 0.9], [0.5, 0.3, 0.3]])
    assert np.allclose(invariants_point(G.T), invariants_point(G))
    assert not np.isclose(invariants_point(G)[0], invariants_point(G + np.eye(3))[0])
    assert not np.isclose(invariants_point(G)[2], invariants_point(G + np.eye(3))[2])


def np_spectra(state, N, M):
    """numpy restatement: compact state [3][nk2][nk1][nk3] -> five shell spectra."""
    nk1 = nk2 = 2 * (N // 2) + 1
    nk3 = N // 2 + 1
    u = state.reshape(3, nk2, nk1, nk3)
    mz = M // 2 + 1
    wrap = lambda i, n: np.where(i <= n // 2, i, i - n)
    k1 = wrap(np.arange(nk1), nk1)
    k2 = wrap(np.arange(nk2), nk2)
    full = np.zeros((3, M, M, mz), complex)  # [c][ky][kx][kz]
    full[:, (k2 % M)[:, None, None], (k1 % M)[None, :, None], np.arange(nk3)[None, None, :]] = u
    ky = np.fft.fftfreq(M, 1.0 / M)[:, None, None]
    kx = np.fft.fftfreq(M, 1.0 / M)[None, :, None]
    kz = np.arange(mz)[None, None, :]
    P = M ** 3
    G = np.zeros((M, M, M, 3, 3))  # [y][x][z][i][j], d u_i / d x_j
    for c in range(3):
        for d, kd in enumerate((kx, ky, kz)):
            G[..., c, d] = np.fft.irfftn(1j * kd * full[c], s=(M, M, M), axes=(0, 1, 2)) * P
    inv = np_invariants(G)
    KX, KY, KZ = np.meshgrid(np.fft.fftfreq(M, 1.0 / M), np.fft.fftfreq(M, 1.0 / M), np.arange(mz), indexing="xy")
    # meshgrid "xy" gives arrays [ky][kx][kz]
    k2 = (KX ** 2 + KY ** 2 + KZ ** 2).astype(np.int64)
    inD = (np.abs(KX) <= N // 2) & (np.abs(KY) <= N // 2) & (KZ <= N // 2)
    wgt = np.where((KZ != 0) & ~((M % 2 == 0) & (KZ == M // 2)), 2.0, 1.0)
    m = np.floor(np.sqrt(k2)).astype(np.int64)
    m = np.where((m + 1) ** 2 <= k2, m + 1, m)
    K = np.where(k2 > m * m + m, m + 1, m)
    nb = int(np.ceil(np.sqrt(3.0 * (N // 2) ** 2))) + 1
    out = np.zeros((5, nb))
    for f in range(5):
        Ik = np.fft.rfftn(inv[..., f], axes=(0, 1, 2)) / P
        val = (np.abs(Ik) ** 2 * wgt)[inD]
        np.add.at(out[f], K[inD], val)
    return out


@pytest.mark.parametrize("kw", [dict(Nx=16, nu=1e-2, C_At=0.5), dict(Nx=16, nu=1e-2, C_At=0.5, C_Smag=0.16)])
def test_spectra_against_numpy(kw):
    o = Oracle(**kw)
    rng = np.random.default_rng(5)
    st = (rng.standard_normal((3, o.nmodes)) + 1j * rng.standard_normal((3, o.nmodes))) * 0.05
    o.set_state(st)
    o.finish_init(False)
    st0 = o.get_state()
    ref = np_spectra(st0, 16, 24)
    got = o.invariant_spectra()
    assert got.shape == ref.shape
    for f in range(5):
        assert np.linalg.norm(got[f] - ref[f]) <= 1e-12 * np.linalg.norm(ref[f]), f
    # after steps: the spectra of the state at the start of the last step
    o.keep_start()
    o.step(2)
    before = o.get_state()
    o.step(1)
    ref = np_spectra(before, 16, 24)
    got = o.invariant_spectra()
    for f in range(5):
        assert np.linalg.norm(got[f] - ref[f]) <= 1e-12 * np.linalg.norm(ref[f]), f
    cur = np_spectra(o.get_state(), 16, 24)
    assert np.linalg.norm(got[0] - cur[0]) > 1e-9 * np.linalg.norm(cur[0])
