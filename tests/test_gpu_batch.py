"""Batched stepping (spns_batch_*): several runs of one shape class advanced by
one kernel launch per pass.  Each slot must match the same run stepped alone
bit for bit (same kernels, same grids, so the same arithmetic and reduction
order), with different physics per slot and different step counts per
chunk."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(**kw):
    import spns_amd
    return spns_amd.Solver(**kw)


def _start(seed, **kw):
    s = _solver(**kw)
    s.init_random(seed=seed, k0=4.0)
    s.finish_init(False)
    return s


def _same(a, b, what):
    sa, sb = a.scalars(), b.scalars()
    assert np.array_equal(a.get_state(), b.get_state()), what
    assert np.array_equal(a.get_rhist(), b.get_rhist()), what
    assert np.array_equal(a.spectrum(), b.spectrum()), what
    for k in ("time", "dt", "ek", "om", "skewness", "iter"):
        assert sa[k] == sb[k], (what, k, sa[k], sb[k])


CLASSES = {
    # DNS, FP32-split, 32^3 (FFT grid 48^3): CFL and fixed steps, rotation
    "dns32_fp32": (dict(Nx=32, precision=1), [dict(nu=1e-3, C_At=0.5), dict(nu=3e-3, C_At=0.5, omega=(0.0, 0.0, 5.0)),
                                              dict(nu=2e-3, fixed_dt=2e-3), dict(nu=5e-4, C_At=0.4, omega=(0.0, 0.0, 1.0))]),
    # LES, FP64, 64^3 (96^3): different C_s and viscosities
    "les64_fp64": (dict(Nx=64, precision=0), [dict(nu=1e-4, C_Smag=0.17, C_At=0.5), dict(nu=1e-3, C_Smag=0.1, C_At=0.5),
                                              dict(nu=3e-4, C_Smag=0.2, C_At=0.5, omega=(0.0, 0.0, 1.0))]),
    # DNS, FP64, 32^3: the fused kernel with the next step's inverse y pass in FP64
    "dns32_fp64": (dict(Nx=32, precision=0), [dict(nu=1e-3, C_At=0.5, force_eps=0.1), dict(nu=2e-3, C_At=0.5),
                                              dict(nu=5e-4, fixed_dt=1e-3, omega=(0.0, 0.0, 2.0))]),
    # the rectangular serving grid 96x64x64 (144x96x96), DNS FP32-split, forced
    "rect96_fp32": (dict(Nx=96, Ny=64, Nz=64, precision=1), [dict(nu=1e-3, C_At=0.5, force_eps=0.1),
                                                             dict(nu=4e-3, C_At=0.5)]),
}


@pytest.mark.parametrize("name", sorted(CLASSES))
def test_batch_matches_alone(name):
    import spns_amd
    shape, runs = CLASSES[name]
    nslots, chunk = len(runs) + 1, 6
    b = spns_amd.Batch(nslots, chunk)
    batched, alone = [], []
    for i, kw in enumerate(runs):
        args = dict(shape, **kw)
        s = _start(10 + i, **args)
        b.attach(i + 1, s)  # slot 0 stays empty
        batched.append(s)
        alone.append(_start(10 + i, **args))
    plans = [[0, 6, 3, 5, 1][:nslots], [0, 2, 6, 0, 6][:nslots], [0, 6, 6, 6, 6][:nslots]]
    for plan in plans:
        b.step(plan)
        b.synchronize()
        for i, s in enumerate(alone):
            if plan[i + 1]:
                s.step(plan[i + 1])
    for i in range(len(runs)):
        _same(batched[i], alone[i], "%s slot %d" % (name, i + 1))
        assert batched[i].query()  # finite
    # detach one, attach a new run in its slot, keep going
    b.detach(2)
    late = _start(99, **dict(shape, **runs[0]))
    late_alone = _start(99, **dict(shape, **runs[0]))
    b.attach(2, late)
    b.step([0, 4, 4] + [4] * (nslots - 3))
    b.synchronize()
    late_alone.step(4)
    _same(late, late_alone, name + " late slot")
    # a run that leaves the batch keeps stepping alone: the batched fused
    # kernel has already run its next inverse y pass (DNS classes), and the
    # run's own step must take that up (slot 1 took 4 steps in the last chunk)
    alone[0].step(4)
    _same(batched[0], alone[0], name + " slot 1")
    b.detach(1)
    batched[0].step(3)
    alone[0].step(3)
    _same(batched[0], alone[0], name + " slot 1 after leaving the batch")
    b.close()
    for s in batched + alone + [late, late_alone]:
        s.close()


def test_batch_rejects_mixed_classes_and_unbatched_lengths():
    import spns_amd
    b = spns_amd.Batch(4, 5)
    a = _start(1, Nx=32, precision=1, nu=1e-3, C_At=0.5)
    b.attach(0, a)
    other = [_start(2, Nx=64, precision=1, nu=1e-3, C_At=0.5),           # another grid
             _start(3, Nx=32, precision=0, nu=1e-3, C_At=0.5),           # another precision
             _start(4, Nx=32, precision=1, nu=1e-3, C_Smag=0.17, C_At=0.5)]  # LES in a DNS class
    for s in other:
        with pytest.raises(spns_amd.SpnsError) as e:
            b.attach(1, s)
        assert e.value.code == spns_amd.SPNS_EINVAL
    b2 = spns_amd.Batch(2, 5)
    odd = _start(5, Nx=16, precision=1, nu=1e-3, C_At=0.5)  # FFT grid 24^3: no batched kernels
    with pytest.raises(spns_amd.SpnsError) as e:
        b2.attach(0, odd)
    assert e.value.code == spns_amd.SPNS_EUNSUPPORTED
    with pytest.raises(spns_amd.SpnsError):
        b.step([6, 0, 0, 0])  # above the chunk
    for x in [b, b2, a, odd] + other:
        x.close()


def test_batch_isolates_a_diverging_slot():
    """A slot whose fixed time step is far above its CFL limit overflows; the
    other slots of the same launches stay bit-identical to their runs alone,
    and spns_query reports the diverged one."""
    import spns_amd
    b = spns_amd.Batch(3, 10)
    good = [_start(7, Nx=32, precision=1, nu=1e-3, C_At=0.5), _start(8, Nx=32, precision=1, nu=2e-3, C_At=0.5)]
    bad = _start(9, Nx=32, precision=1, nu=1e-4, fixed_dt=5.0)
    b.attach(0, good[0])
    b.attach(1, bad)
    b.attach(2, good[1])
    for _ in range(4):
        b.step([10, 10, 10])
        b.synchronize()
    with pytest.raises(spns_amd.SpnsError) as e:
        bad.query()
    assert e.value.code == spns_amd.SPNS_ENONFINITE
    for i, g in enumerate(good):
        assert g.query()
        ref = _start(7 + i, Nx=32, precision=1, nu=(1e-3, 2e-3)[i], C_At=0.5)
        ref.step(40)
        _same(g, ref, "good slot %d" % i)
        ref.close()
    for x in [b, bad] + good:
        x.close()


def test_batch_status_and_fetch_without_device_reads():
    """spns_batch_slot_scalars (the copy the batch makes after each chunk) and
    spns_batch_fetch into pinned host arrays give exactly what the context's
    own synchronous reads give."""
    import spns_amd
    b = spns_amd.Batch(3, 8)
    runs = [_start(20 + i, Nx=64, precision=1, nu=1e-3 * (i + 1), C_At=0.5) for i in range(2)]
    b.attach(0, runs[0])
    b.attach(2, runs[1])
    b.step([8, 0, 5])
    b.synchronize()
    st = spns_amd.Pinned((3, runs[1].nmodes), np.complex128)
    ek = spns_amd.Pinned((runs[1].nbins,))
    b.fetch(2, st.array, ek.array)
    b.synchronize()
    for slot, s in ((0, runs[0]), (2, runs[1])):
        fast, ref = b.slot_scalars(slot), s.scalars()
        for k in ("time", "dt", "ek", "om", "skewness", "maxvel", "iter"):
            assert fast[k] == ref[k], (slot, k)
    assert np.array_equal(st.array, runs[1].get_state())
    assert np.array_equal(ek.array, runs[1].spectrum())
    st.close()
    ek.close()
    with pytest.raises(spns_amd.SpnsError):
        b.slot_scalars(1)  # empty slot
    b.close()
    for s in runs:
        s.close()
