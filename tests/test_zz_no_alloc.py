"""spns_invariant_spectra_noalloc (the entry point that allocates nothing:
its partial sums live in a fixed device array) against spns_invariant_spectra,
through the C ABI: bit-identical spectra on two calls in a row on one GPU,
the second on a smaller grid (the array keeps the first call's contents),
and a clear error for a grid whose sums do not fit the array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(N, steps, prec=0):
    import spns_amd
    s = spns_amd.Solver(Nx=N, nu=2e-3, C_At=0.5, precision=prec, gradient_invariants=True)
    s.init_random(seed=N, k0=4.0)
    s.finish_init(False)
    s.step(steps)
    return s


def _spectra(s, fn):
    out = np.zeros((5, s.nbins))
    rc = getattr(s.L, fn)(s.h, s.nbins, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


def test_noalloc_matches_allocating_entry_point():
    import spns_amd
    for N, prec in ((64, 0), (32, 1)):  # the second call: a smaller grid (and FP32-split)
        s = _run(N, 3, prec)
        rc0, ref = _spectra(s, "spns_invariant_spectra")
        rc1, out = _spectra(s, "spns_invariant_spectra_noalloc")
        assert rc0 == 0 and rc1 == 0, spns_amd.load_library().spns_last_error(s.h)
        assert np.array_equal(out, ref), N
        assert np.all(np.isfinite(out)) and out[0].sum() != 0.0
        s.close()


def test_noalloc_rejects_grids_beyond_its_array():
    import spns_amd
    s = _run(128, 1, 1)  # 112 shells x 1025 blocks x 5 > 2^19 doubles
    rc, _ = _spectra(s, "spns_invariant_spectra_noalloc")
    assert rc == spns_amd.SPNS_EINVAL
    assert b"fixed scratch" in spns_amd.load_library().spns_last_error(s.h)
    rc0, ref = _spectra(s, "spns_invariant_spectra")  # the allocating one still serves it
    assert rc0 == 0 and np.all(np.isfinite(ref))
    s.close()
