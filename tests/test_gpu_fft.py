"""FFT_package drop-in (src/FFTW_package.cpp:7-42) on the GPU against numpy.

Layouts follow the reference: complex [y][x][Mz/2+1], real [x][y][2(Mz/2+1)].
Tolerances: FP64 1e-13 relative to the field norm; FP32 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.mark.parametrize("shape", [(48, 48, 48), (32, 48, 24), (96, 72, 144), (288, 144, 96)])
@pytest.mark.parametrize("prec", [0, 1])
def test_r2c_c2r_vs_numpy(shape, prec):
    import spns_amd

    dev = _dev()
    Mx, My, Mz = shape
    nk = Mz // 2 + 1
    rng = np.random.default_rng(sum(shape) + prec)
    r = rng.standard_normal(shape)
    rp = np.zeros((Mx, My, 2 * nk))
    rp[:, :, :Mz] = r
    dt = torch.float64 if prec == 0 else torch.float32
    rin = torch.from_numpy(rp).to(dev, dt)
    cout = torch.zeros((My, Mx, nk, 2), dtype=dt, device=dev)
    f = spns_amd.Fft(Mx, My, Mz, precision=prec)
    torch.cuda.synchronize()  # inputs and outputs were filled on torch's stream
    f.r2c(rin.data_ptr(), cout.data_ptr())
    f.synchronize()
    c = cout.double().cpu().numpy()
    c = c[..., 0] + 1j * c[..., 1]
    ref = np.fft.rfftn(r).transpose(1, 0, 2)
    tol = 1e-13 if prec == 0 else 2e-6
    assert np.linalg.norm(c - ref) / np.linalg.norm(ref) < tol
    # c2r back (input destroyed, as in the reference)
    rout = torch.zeros((Mx, My, 2 * nk), dtype=dt, device=dev)
    torch.cuda.synchronize()  # inputs and outputs were filled on torch's stream
    f.c2r(cout.data_ptr(), rout.data_ptr())
    f.synchronize()
    back = rout.double().cpu().numpy()[:, :, :Mz] / r.size
    assert np.linalg.norm(back - r) / np.linalg.norm(r) < tol
    f.close()


def test_c2r_single_mode_closed_form():
    """A single mode (k1,k2,k3) = (2,-3,5) with coefficient 1 gives
    2 cos(k.x) on the grid (c2r is unnormalised)."""
    import spns_amd

    dev = _dev()
    Mx = My = Mz = 48
    nk = Mz // 2 + 1
    c = np.zeros((My, Mx, nk), dtype=complex)
    c[(-3) % My, 2, 5] = 1.0
    ct = torch.from_numpy(np.stack([c.real, c.imag], -1)).to(dev)
    rout = torch.zeros((Mx, My, 2 * nk), dtype=torch.float64, device=dev)
    f = spns_amd.Fft(Mx, My, Mz)
    torch.cuda.synchronize()  # inputs and outputs were filled on torch's stream
    f.c2r(ct.data_ptr(), rout.data_ptr())
    f.synchronize()
    x = np.arange(Mx)[:, None, None] * 2 * np.pi / Mx
    y = np.arange(My)[None, :, None] * 2 * np.pi / My
    z = np.arange(Mz)[None, None, :] * 2 * np.pi / Mz
    ref = 2 * np.cos(2 * x - 3 * y + 5 * z)
    assert np.abs(rout.cpu().numpy()[:, :, :Mz] - ref).max() < 1e-13
