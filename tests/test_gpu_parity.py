"""Parity of the HIP step (FP64 plain form) with the FP64 oracle.

Same seeded initial field in both; after N steps the velocity state, the
R history, E(k), time, dt, energy and enstrophy must agree to relative L2
<= 1e-12 (tolerance stated by the task for the plain FP64 form).  The QA
replays check the product against the reference's own printed log rows.
"""
import numpy as np
import pytest

import qa_driver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _solver(**kw):
    import spns_amd
    return spns_amd.Solver(**kw)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


INIT_TOL = 1e-14


def pair(seed=1, k0=4.0, **kw):
    """A GPU solver and an oracle, each generating the same seeded field on
    its own side (GPU kernel / host C); the two fields agree to 1e-14
    (device vs host libm in Box-Muller, FP64 energy sums in the scaling)."""
    g = _solver(**kw)
    for k in ("use_graph", "precision", "gradient_invariants"):
        kw.pop(k, None)
    o = Oracle(**kw)
    g.init_random(seed=seed, k0=k0)
    o.init_random(seed=seed, k0=k0)
    e = rel(g.get_state(), o.get_state())
    assert e <= INIT_TOL, ("initial field", e)
    return g, o


def compare(g, o, what=""):
    sg, so = g.scalars(), o.scalars()
    out = {
        "state": rel(g.get_state(), o.get_state()),
        "rhist": rel(g.get_rhist(), o.get_rhist()),
        "ek_bins": rel(g.spectrum(), o.spectrum()),
        "time": rel(sg["time"], so["time"]),
        "dt": rel(sg["dt"], so["dt"]),
        "ek": rel(sg["ek"], o.energy()),
        "om": rel(sg["om"], o.enstrophy()),
    }
    print(what, {k: "%.2e" % v for k, v in out.items()}, flush=True)
    for k, v in out.items():
        assert v <= TOL, (what, k, v)
    return out


CASES = {
    "dns_fixed_dt": dict(Nx=32, nu=0.01, fixed_dt=1e-3),
    "dns_cfl": dict(Nx=32, nu=1e-3, C_At=0.5),
    "les": dict(Nx=32, nu=1e-4, C_Smag=0.17, C_At=0.5),
    "rotation": dict(Nx=32, nu=1e-3, C_At=0.5, omega=(1.5, -2.1, 3.0)),
    "const_power": dict(Nx=32, nu=2e-3, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2),
    "self_adaptive": dict(Nx=32, nu=1e-3, C_At=0.5, isSelfAdaptive=True),
    "rect_grid": dict(Nx=64, Ny=32, Nz=48, nu=2e-3, C_At=0.5),
    "les_rot_64": dict(Nx=64, nu=5e-4, C_Smag=0.16, C_At=0.5, omega=(0.0, 0.0, 5.0)),
    # oversampled z axis: z transforms of lengths 512, 768, 1536 with few retained k3
    "zwave512": dict(Nx=32, Mz=512, nu=1e-3, C_At=0.5),
    "zwave768_forced": dict(Nx=32, Mz=768, nu=2e-3, C_At=0.5, force_eps=0.1),
    "zwave1536_sa": dict(Nx=32, Mz=1536, nu=1e-3, C_At=0.5, isSelfAdaptive=True),
    # the z transform of the 512^3 run (768 points, 257 retained k3) on a
    # small x-y cross-section, DNS and LES
    "z768_full_k3": dict(Nx=16, Ny=16, Nz=512, nu=2e-3, C_At=0.5, force_eps=0.1),
    "z768_full_k3_les": dict(Nx=16, Ny=16, Nz=512, nu=1e-4, C_Smag=0.17, C_At=0.5),
    "z1536_full_k3": dict(Nx=16, Ny=16, Nz=1024, nu=2e-3, C_At=0.5),
    # a 1:1:2 box: length factors (2, 2, 1) after the lcm expansion of
    # src/MAIN.cpp:177-226, only modes with k_i divisible by L_i retained
    "noncubic_112": dict(Nx=32, Ny=32, Nz=16, Mx=48, My=48, Mz=24, Lx=2, Ly=2, Lz=1, isNotCubic=True, nu=1e-3, C_At=0.5),
    # the y transform of the 1024^3 run (1536 points, 1025 retained k2) on
    # a thin x-z cross-section: the fused y pass + epilogue at that length
    "y1536_full_k2": dict(Nx=16, Ny=1024, Nz=16, nu=2e-3, C_At=0.5, force_eps=0.1),
    # the x transform of the 512^3 run (768 points, 513 retained k1): the FP64
    # inverse x pass runs both transforms of each dual job as one
    "x768_full_k1": dict(Nx=512, Ny=16, Nz=16, nu=2e-3, C_At=0.5),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("graph", [False, True])
def test_step_parity(name, graph):
    kw = dict(CASES[name])
    g, o = pair(seed=7, use_graph=graph, **kw)
    g.finish_init(False)
    o.finish_init(False)
    compare(g, o, name + " init")
    for chunk in (1, 4, 15):
        g.step(chunk)
        o.step(chunk)
        compare(g, o, "%s step+%d" % (name, chunk))


def test_decaying_dns_32_full_200_steps():
    """32^3 decaying DNS, nu=0.01, fixed dt=1e-3, 200 steps (seed 1, k0=4)."""
    g, o = pair(seed=1, k0=4.0, Nx=32, nu=0.01, fixed_dt=1e-3)
    g.finish_init(False)
    o.finish_init(False)
    ek_hist_g, ek_hist_o = [], []
    for _ in range(200):
        g.step(1)
        o.step(1)
        ek_hist_g.append(g.scalars()["ek"])
        ek_hist_o.append(o.energy())
    compare(g, o, "32^3 decaying DNS, 200 steps")
    assert rel(ek_hist_g, ek_hist_o) <= TOL
    o.antitransform()  # physical field of the current state on both sides
    assert rel(g.velocity_physical(), o.velocity_physical()) <= TOL


def test_velocity_physical_matches():
    g, o = pair(seed=3, Nx=32, nu=0.01)
    g.finish_init(False)
    o.finish_init(False)
    assert rel(g.velocity_physical(), o.velocity_physical()) < TOL


def test_skewness_matches_oracle():
    g, o = pair(seed=5, Nx=32, nu=0.01, fixed_dt=1e-3)
    g.finish_init(False)
    o.finish_init(False)
    g.step(3)
    o.step(3)
    so = o.scalars()
    P = 48 ** 3
    sk = (so["skew_s3"] / P) / (so["skew_s2"] / P) ** 1.5
    assert abs(g.scalars()["skewness"] - sk) <= 1e-12 * abs(sk)


FIX = qa_driver.load_fixtures()


@pytest.mark.parametrize("name", sorted(FIX))
def test_qa_replay_on_gpu(name):
    """Every per-step row of the reference QA log, reproduced by the GPU."""
    fx = FIX[name]
    rows, _ = qa_driver.run(fx["params"], lambda **kw: _solver(**kw))
    ref = fx["rows"]
    assert len(rows) == len(ref)
    bad = [(i, r, e) for i, (r, e) in enumerate(zip(rows, ref)) if r != e]
    assert len(bad) == 0, bad[:3]


@pytest.mark.parametrize("kw", [dict(Nx=32, nu=1e-3, C_At=0.5), dict(Nx=32, nu=1e-4, C_Smag=0.17, C_At=0.5),
                                dict(Nx=32, nu=1e-3, C_At=0.5, isSelfAdaptive=True)])
def test_velocity_step_start(kw):
    """Between steps the physical field of the last step's start state is the
    field of the state before that step, bit for bit, and fetching it leaves
    the CFL step alone (Recalculate_TimeStep gives the same dt)."""
    import spns_amd
    g = _solver(keep_step_start=True, **kw)
    g.init_random(seed=4, k0=4.0)
    g.finish_init(False)
    assert np.array_equal(g.velocity_physical_step_start(), g.velocity_physical())
    for _ in range(3):
        before = g.velocity_physical()
        g.step(1)
        g.recalc_timestep()
        dt = g.scalars()["dt"]
        got = g.velocity_physical_step_start()
        g.recalc_timestep()
        assert g.scalars()["dt"] == dt
        assert np.array_equal(got, before)
    h = _solver(**kw)
    h.init_random(seed=4, k0=4.0)
    h.finish_init(False)
    h.step(1)
    if not kw.get("isSelfAdaptive"):
        with pytest.raises(spns_amd.SpnsError):
            h.velocity_physical_step_start()
    g.close()
    h.close()


@pytest.mark.parametrize("kw", [dict(Nx=32, nu=1e-3, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2),
                                dict(Nx=32, nu=1e-3, C_At=0.5, isSelfAdaptive=True),
                                dict(Nx=16, Ny=16, Nz=512, nu=2e-3, C_At=0.5, omega=(0.0, 0.0, 3.0)),
                                dict(Nx=32, Ny=32, Nz=16, Mx=48, My=48, Mz=24, Lx=2, Ly=2, Lz=1, isNotCubic=True, nu=1e-3, C_At=0.5),
                                dict(Nx=16, Ny=128, Nz=16, nu=2e-3, C_At=0.5),
                                dict(Nx=16, Ny=256, Nz=24, nu=2e-3, C_At=0.5, force_eps=0.1),
                                dict(Nx=16, Ny=1024, Nz=16, nu=2e-3, C_At=0.5, force_eps=0.1)])
@pytest.mark.parametrize("precision", [0, 1])
def test_fused_epilogue_matches_split(monkeypatch, kw, precision):
    """The DNS step's forward y pass fused with the epilogue against the two
    separate kernels (SPNS_EPI=split): same state, history and scalars up to
    the summation order of the shell sums (which feeds the forcing and dt)."""
    out = {}
    for mode in ("split", "fused"):
        if mode == "split":
            monkeypatch.setenv("SPNS_EPI", "split")
        else:
            monkeypatch.delenv("SPNS_EPI", raising=False)
        g = _solver(precision=precision, **kw)
        assert (g.pass_kernels()[1] > 0) == (mode == "fused")  # the fused kernel really runs (or really does not)
        g.init_random(seed=9, k0=4.0)
        g.finish_init(False)
        g.step(6)
        out[mode] = (g.get_state(), g.get_rhist(), g.spectrum(), g.scalars())
        g.close()
    s, f = out["split"], out["fused"]
    tol = 1e-12 if precision == 0 else 1e-6
    assert rel(f[0], s[0]) <= tol
    assert rel(f[1], s[1]) <= tol
    assert rel(f[2], s[2]) <= tol
    for k in ("time", "dt", "ek", "om", "skewness"):
        assert abs(f[3][k] - s[3][k]) <= tol * max(1.0, abs(s[3][k])), k


YINV_CASES = [dict(Nx=32, nu=1e-3, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2),
              dict(Nx=32, nu=1e-3, C_At=0.5, isSelfAdaptive=True),
              dict(Nx=64, Ny=32, Nz=48, nu=2e-3, C_At=0.5, omega=(0.0, 0.0, 3.0)),
              dict(Nx=16, Ny=128, Nz=16, nu=2e-3, C_At=0.5),
              dict(Nx=16, Ny=256, Nz=24, nu=2e-3, C_At=0.5, force_eps=0.1),
              dict(Nx=16, Ny=512, Nz=16, nu=2e-3, C_At=0.5),
              dict(Nx=16, Ny=1024, Nz=16, nu=2e-3, C_At=0.5, force_eps=0.1)]


@pytest.mark.parametrize("kw", YINV_CASES)
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("graph", [False, True])
def test_fused_inverse_y_matches_separate(monkeypatch, kw, precision, graph):
    """The next step's inverse y pass inside the fused forward y pass +
    epilogue kernel against the separate pass (SPNS_YINV=0): bit for bit,
    including across the calls that rewrite X1 or change the state between
    steps (velocity passes, the step-start field, set_state)."""
    out = {}
    for mode in ("separate", "fused"):
        if mode == "separate":
            monkeypatch.setenv("SPNS_YINV", "0")
        else:
            monkeypatch.delenv("SPNS_YINV", raising=False)
        g = _solver(precision=precision, use_graph=graph, keep_step_start=not kw.get("isSelfAdaptive"), **kw)
        assert g.pass_kernels()[1] == (2 if mode == "fused" else 1)
        g.init_random(seed=11, k0=4.0)
        g.finish_init(False)
        g.step(3)
        v1 = g.velocity_physical()
        g.step(2)
        v2 = g.velocity_physical_step_start() if not kw.get("isSelfAdaptive") else v1
        g.step(1)
        g.set_state(g.get_state())
        g.step(4)
        out[mode] = (g.get_state(), g.get_rhist(), g.spectrum(), g.scalars(), v1, v2)
        g.close()
    s, f = out["separate"], out["fused"]
    for a, b in ((s[0], f[0]), (s[1], f[1]), (s[2], f[2]), (s[4], f[4]), (s[5], f[5])):
        assert np.array_equal(a, b)
    for k in ("time", "dt", "ek", "om", "skewness", "iter"):
        assert s[3][k] == f[3][k], k


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_inverse_y_across_state_changes(monkeypatch, precision):
    """Every call between steps that changes the state or rewrites the work
    buffers (conjugate correction, the QA energy check, invariant spectra, a
    forcing spectrum, an antitransform, a reconfigured run, the K41 and
    physical-field initialisations) leaves the fused next-step inverse y pass
    bit-identical to the separate pass."""
    kw = dict(Nx=32, nu=1e-3, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2)
    out = {}
    for mode in ("separate", "fused"):
        if mode == "separate":
            monkeypatch.setenv("SPNS_YINV", "0")
        else:
            monkeypatch.delenv("SPNS_YINV", raising=False)
        g = _solver(precision=precision, gradient_invariants=True, keep_step_start=True, **kw)
        assert g.pass_kernels()[1] == (2 if mode == "fused" else 1)
        g.init_random(seed=13, k0=4.0)
        g.finish_init(False)
        rec = []
        g.step(2)
        g.ccc()
        g.step(2)
        g.check_energy(atol=1e-3)
        g.step(1)
        rec.append(g.invariant_spectra())
        g.step(2)
        ek = g.spectrum()
        g.forcing_energy_cascade(ek, len(ek) - 1)
        g.step(2)
        g.antitransform()
        g.step(2)
        rec.append(g.get_state())
        g.reconfigure(nu=2e-3)
        g.init_random(seed=14, k0=4.0)
        g.finish_init(False)
        g.step(3)
        rec.append(g.get_state())
        g.init_k41(4)
        g.finish_init(False)
        g.step(2)
        rec.append(g.get_state())
        u = g.velocity_physical()
        g.init_from_physical(u[0], u[1], u[2])
        g.finish_init(False)
        g.step(2)
        rec.append(g.get_state())
        rec.append(g.get_rhist())
        out[mode] = rec
        g.close()
    for a, b in zip(out["separate"], out["fused"]):
        assert np.array_equal(a, b)
