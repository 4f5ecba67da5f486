"""Velocity-gradient invariant spectra on the GPU (QG, QS, RG, RS, V2;
src/HIT_LES.cpp:1297-1319, src/tensor_algebra.cpp:250-303) against the
oracle, whose spectra are pinned to numpy in test_oracle_invariants.py.

Same seeded field on both sides; spectra before any step (of the current
field) and after steps (of the field at the start of the last step, the
gradients the reference's last convective term left).  Stepping resumes
after a spectra call and stays at step parity: the spectra reuse the step's
scratch buffers only.  FP64 plain form within relative L2 1e-12 per
invariant; the FP32-split form within 5e-6 (products of up to four single-
precision gradients; about 5e-7 measured)."""
import numpy as np
import pytest

from test_gpu_parity import compare, pair

pytestmark = pytest.mark.gpu

CASES = {
    "dns32": dict(Nx=32, nu=1e-3, C_At=0.5),
    "les32": dict(Nx=32, nu=1e-4, C_Smag=0.17, C_At=0.5),
    "rect": dict(Nx=32, Ny=16, Nz=24, nu=2e-3, C_At=0.5),
    "sa32": dict(Nx=32, nu=1e-3, C_At=0.5, isSelfAdaptive=True),
    "z768_full_k3": dict(Nx=16, Ny=16, Nz=512, nu=2e-3, C_At=0.5, force_eps=0.1),
    "noncubic_112": dict(Nx=32, Ny=32, Nz=16, Mx=48, My=48, Mz=24, Lx=2, Ly=2, Lz=1, isNotCubic=True, nu=1e-3, C_At=0.5),
}
NAMES = ("QG", "QS", "RG", "RS", "V2")


def check(got, ref, tol, what):
    assert got.shape[0] == 5
    nb = min(got.shape[1], ref.shape[1])
    # bins past the oracle's last shell stay zero on the GPU
    assert not np.any(got[:, nb:])
    errs = {}
    for f in range(5):
        r = ref[f, :nb]
        errs[NAMES[f]] = float(np.linalg.norm(got[f, :nb] - r) / np.linalg.norm(r))
    print(what, {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, (what, k, v)


@pytest.mark.parametrize("name", sorted(CASES))
def test_invariant_spectra_fp64(name):
    g, o = pair(seed=11, gradient_invariants=True, **CASES[name])
    g.finish_init(False)
    o.finish_init(False)
    o.keep_start()
    check(g.invariant_spectra(), o.invariant_spectra(), 1e-12, name + " init")
    for chunk in (1, 3):
        g.step(chunk)
        o.step(chunk)
        check(g.invariant_spectra(), o.invariant_spectra(), 1e-12, "%s step+%d" % (name, chunk))
    g.step(2)
    o.step(2)
    compare(g, o, name + " after spectra")
    g.close()


@pytest.mark.parametrize("name", ["dns32", "les32", "z768_full_k3"])
def test_invariant_spectra_fp32_split(name):
    import spns_amd
    g, o = pair(seed=11, gradient_invariants=True, precision=spns_amd.PREC_FP32_SPLIT, **CASES[name])
    g.finish_init(False)
    o.finish_init(False)
    o.keep_start()
    check(g.invariant_spectra(), o.invariant_spectra(), 5e-6, name + " fp32 init")
    g.step(2)
    o.step(2)
    check(g.invariant_spectra(), o.invariant_spectra(), 5e-6, name + " fp32 step+2")
    g.close()


def test_invariant_spectra_needs_the_flag():
    import spns_amd
    g = spns_amd.Solver(Nx=16, nu=1e-2, C_At=0.5)
    g.init_random(seed=1, k0=4.0)
    g.finish_init(False)
    with pytest.raises(spns_amd.SpnsError):
        g.invariant_spectra()
    g.close()
