"""Benchmark: RK (AB2) steps/s and HBM roofline at 512^3, FP32-split vs FP64,
plus the ensemble serving mix.

Workload (the 512^3 forced DNS): Nx=Ny=Nz=512 retained modes, FFT grid
768^3, nu=2e-4, constant-power forcing eps=0.1 on shells 1<=K<=2, CFL step
with C_At=0.5, seeded random solenoidal field with
E(k) ~ k^4 exp(-2 (k/k0)^2), k0=4, u_rms=1 (synthetic data).  --N 1024 runs
the 1024^3 case (FFT grid 1536^3, nu=8e-5).

Timed region: K captured steps (hipGraph) with the state resident in HBM,
bracketed by stream and device synchronisation.  `value` is the FP32-split
form's steps/s; the FP64 plain form is measured in the same run.

The roofline object covers the dominant kernel of the FP32-split step:
achieved = its algorithmic bytes per launch / its mean launch time, the
time from HIP events on the library's stream (spns_profile_passes).
`traffic` comes from the committed rocprofv3 PMC summary of the same
kernel and workload when present (profiles/*.json), else null.
`copy_peak` is a measured streaming copy (a hand-written kernel), the
practical ceiling beside 8 TB/s.

cpu_baseline: the FP64 C oracle (a port of the reference step) on the
host's cores: one step of the same forced-DNS setup at the benchmark size
(all field pages touched before the timed step), the bounded sample.  A
smaller --cpu-N is scaled to the benchmark size by P log2 P and says so.

ensemble: the serving mix (Poisson arrivals at 20/s of mixed runs), its
throughput and p50/p99 completion latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--N 512]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # MI355X HBM3E bytes/s
METRIC = "RK steps/s + HBM % at 512^3, FP32-split vs FP64; ensemble p50/p99 latency"


def dims(N):
    M = (3 * N) // 2
    nk1 = nk2 = 2 * (N // 2) + 1
    nk3 = N // 2 + 1
    return M, nk1, nk2, nk3


def pass_bytes(N, prec, les=False, fused_epilogue=0):
    """Algorithmic (compulsory) HBM bytes of each pass of one step.  With the
    DNS forward y pass fused with the epilogue (fused_epilogue >= 1: one
    kernel, timed as fwd_y) the convective spectra never reach HBM; with the
    next step's inverse y pass in that kernel too (2) the step's inverse y
    pass moves nothing and the new state is not read back from HBM (its
    outputs are written by the fused kernel).  The state is FP64 (16 bytes a
    mode) and the R history is held in the transform type (sc)."""
    M, nk1, nk2, nk3 = dims(N)
    sc = 16 if prec == 0 else 8  # complex of the transform precision
    KN = nk1 * nk2 * nk3
    x1 = M * nk1 * nk3  # one field after the y pass
    x2 = M * M * nk3    # one field after the x pass
    nf = 9 if les else 3
    b = {
        "inv_y": 3 * KN * 16 + 6 * x1 * sc,
        "inv_x": 6 * x1 * sc + 9 * x2 * sc,
        "zpass": 9 * x2 * sc + nf * x2 * sc,
        "fwd_x": nf * x2 * sc + nf * x1 * sc,
        "fwd_y": nf * x1 * sc + nf * KN * sc,
        "epilogue": nf * KN * sc + 6 * KN * 16 + 6 * KN * sc,
    }
    if fused_epilogue:
        b["fwd_y"] = nf * x1 * sc + 6 * KN * 16 + 6 * KN * sc
        b["epilogue"] = 0
    if fused_epilogue == 2:
        b["fwd_y"] += 6 * x1 * sc
        b["inv_y"] = 0
    return b


def model_bytes(N, prec, les=False):
    """Compulsory-byte model of a step: n_fft * 2 * s_t * M^3 +
    6 * (16 + 2 * s_t) * K_N (one write + read of an M^3 field per 3-D
    transform; the FP64 state and the R history, held in the transform type,
    read and written once)."""
    M, nk1, nk2, nk3 = dims(N)
    n_fft = 21 if les else 15
    s_t = 8 if prec == 0 else 4
    return n_fft * 2 * s_t * M ** 3 + 6 * (16 + 2 * s_t) * nk1 * nk2 * nk3


def sync_all(solver):
    solver.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def workload_nu(N):
    """nu of the forced-DNS workloads: 2e-4 at 512^3, 8e-5 at 1024^3."""
    return 8e-5 if N >= 1024 else 2e-4


LES_NU, LES_CS, LES_K0 = 1e-4, 0.17, 6.0  # the Smagorinsky LES workload (--les)


def run_form(N, prec, steps, warmup, seed, prof_steps, les=False):
    import spns_amd
    if les:
        # decaying Smagorinsky LES, Delta = 2 pi / N, CFL step
        s = spns_amd.Solver(Nx=N, nu=LES_NU, C_Smag=LES_CS, C_At=0.5, delta=2.0 * math.pi / N, precision=prec,
                            use_graph=True)
    else:
        s = spns_amd.Solver(Nx=N, nu=workload_nu(N), C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2,
                            precision=prec, use_graph=True)
    s.init_random(seed=seed, k0=LES_K0 if les else 4.0, urms=1.0)
    s.finish_init(False)
    s.step(warmup)
    sync_all(s)
    t0 = time.perf_counter()
    s.step(steps)
    sync_all(s)
    t1 = time.perf_counter()
    ev_ms = s.last_step_ms()
    sc = s.scalars()
    ek = s.spectrum()
    passes = s.profile_passes(prof_steps) if prof_steps > 0 else {}
    fused = s.pass_kernels()[1]
    s.close()
    return {"wall_s": t1 - t0, "event_ms": ev_ms, "scalars": sc, "ek": ek, "passes": passes, "fused": fused}


def load_traffic(kernel, workload):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "pmc_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("bytes_per_launch")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_cpu, threads, steps, les=False):
    """Median wall time of `steps` steps of the same forced-DNS (or LES)
    setup at N=n_cpu on the host (the FP64 C oracle, 1 process x OpenMP
    threads)."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle.oracle import Oracle, lib
    import numpy as np
    if les:
        o = Oracle(n_cpu, nu=LES_NU, C_Smag=LES_CS, C_At=0.5, delta=2.0 * math.pi / n_cpu)
    else:
        o = Oracle(n_cpu, nu=2e-4, C_At=0.5, force_eps=0.1, force_kmin=1, force_kmax=2)
    rng = np.random.default_rng(0)
    st = (rng.standard_normal((3, o.nmodes)) + 1j * rng.standard_normal((3, o.nmodes))) * 1e-3
    o.set_state(st)
    o.finish_init(False)
    o.prefault()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.step(1)
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2], ts, lib().orc_num_threads()


def run_ensemble(n, seed, batched=True, rate=20.0, chunk=20):
    """The serving mix (Poisson arrivals at `rate` req/s, mixed grids,
    DNS/LES, viscosities and rotation rates, 100-2000 steps each,
    FP32-split), first n requests: throughput and completion-latency
    percentiles.  batched: one launch per pass over the runs of a shape
    class (spns_batch); else one graph per run on its own stream."""
    from spns_amd import serve
    reqs = serve.draw_requests(n, rate=rate, seed=seed, steps=(100, 2000))
    srv = serve.Server(tuner=None if batched else serve.Tuner(), batched=batched, chunk=chunk)
    stats, done = srv.run(reqs)
    if batched:
        stats["policy"] = {"batched": True, "nslots": srv.nslots, "chunk": srv.chunk}
    else:
        stats["policy"] = {"batched": False, "inflight_work": srv.inflight_work, "min_inflight": srv.min_inflight}
    steps = sum(r.steps for r in reqs)
    out = dict(stats)
    out.update({"arrival_rate_per_s": rate, "seed": seed, "total_steps": steps,
                "grids": sorted(set("x".join(map(str, r.grid)) for r in reqs)),
                "les_fraction": sum(r.C_Smag > 0 for r in reqs) / max(1, len(reqs))})
    if not batched:
        out["zpass_variant_choice"] = {"%s" % (k,): v for k, v in srv.tuner.choice.items()}
    return out


def copy_calibration(mib):
    """A plain device copy of known size: the reference reading for the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) of a profiled run."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4, "launches": 3}))


def copy_peak(mib=4096, reps=10):
    """Measured practical HBM ceiling beside the 8 TB/s data-sheet peak: the
    library's streaming copy kernel (16-byte non-temporal loads and stores,
    spns_copy_bandwidth) between two device buffers of mib MiB, HIP events,
    read + written bytes; a torch device copy of the same size beside it."""
    import spns_amd
    out = {"GB_s": spns_amd.copy_bandwidth(mib << 20, reps), "bytes_per_copy": 2 * (mib << 20),
           "how": "hand-written streaming copy kernel (k_stream_copy), %d MiB each way, mean of %d, read+write bytes" % (mib, reps)}
    try:
        import torch
        n = mib * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        out["torch_copy_GB_s"] = 2.0 * n * 4 / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    except Exception:
        pass
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--prof-steps", type=int, default=3)
    ap.add_argument("--cpu-N", type=int, default=0, help="oracle size for cpu_baseline (0: the benchmark's N)")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--ensemble-requests", type=int, default=32,
                    help="requests of the serving mix to run (0: skip)")
    ap.add_argument("--ensemble-rate", type=float, default=20.0, help="Poisson arrival rate of the serving mix (req/s)")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving mix (--ensemble-requests at --ensemble-rate) and print its JSON line")
    ap.add_argument("--chunk", type=int, default=20, help="steps per batched chunk in the serving loop")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,fp32")
    ap.add_argument("--les", action="store_true",
                    help="the Smagorinsky LES workload (nu=1e-4, C_s=0.17, Delta=2pi/N, k0=6) instead of the "
                         "forced DNS; use with --N 256")
    ap.add_argument("--copy-peak", action="store_true", help="only print the measured copy bandwidth and exit")
    ap.add_argument("--no-copy-peak", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.copy_peak:
        print(json.dumps(copy_peak()))
        return
    if a.serving_only:
        out = run_ensemble(a.ensemble_requests, a.seed, rate=a.ensemble_rate, chunk=a.chunk)
        out["metric"] = "serving mix: p50/p99 completion latency at a steady Poisson rate"
        print(json.dumps(out))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: several GPUs would shard the step as a slab decomposition "
              "(one all-to-all per 3-D transform), which is not implemented.  Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    copy = None
    if not a.no_copy_peak:
        # in a child process: torch's HIP runtime and the library's stay apart
        import subprocess
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-peak"], capture_output=True, text=True)
        if r.returncode == 0 and r.stdout.strip():
            copy = json.loads(r.stdout.strip().splitlines()[-1])
            print("bench: device copy %.0f GB/s" % copy["GB_s"], file=sys.stderr, flush=True)
        else:
            print("bench: copy peak not measured: %s" % r.stderr.strip()[-300:], file=sys.stderr, flush=True)

    N = a.N
    M = dims(N)[0]
    res = {}
    forms = [f.strip() for f in a.forms.split(",") if f.strip()]
    for form in forms:
        prec = 0 if form == "fp64" else 1
        print("bench: %s N=%d (M=%d) warmup=%d steps=%d" % (form, N, M, a.warmup, a.steps), file=sys.stderr, flush=True)
        res[form] = run_form(N, prec, a.steps, a.warmup, a.seed, a.prof_steps, les=a.les)
        r = res[form]
        print("bench: %s %.3f ms/step (events %.3f ms over %d steps)" % (form, 1e3 * r["wall_s"] / a.steps,
              r["event_ms"] / a.steps, a.steps), file=sys.stderr, flush=True)

    main_form = "fp32" if "fp32" in res else forms[0]
    r = res[main_form]
    ms = 1e3 * r["wall_s"] / a.steps
    sps = a.steps / r["wall_s"]
    prec_main = 1 if main_form == "fp32" else 0

    # the step's kernels (spns_pass_kernels): 0 separate forward y pass and
    # epilogue, 1 fused, 2 fused with the next step's inverse y pass;
    # spns_profile_passes times the fused kernel under fwd_y
    def fused(rr):
        return rr["fused"]

    # dominant kernel roofline from per-pass HIP-event timings
    pb = pass_bytes(N, prec_main, les=a.les, fused_epilogue=fused(r))
    passes = r["passes"]
    roof = None
    if passes:
        dom = max((k for k in pb if k in passes), key=lambda k: passes[k])
        ach = pb[dom] / (passes[dom] * 1e-3)
        wl = "N%d_%s" % (N, main_form)
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": ach / PEAK_HBM, "traffic": load_traffic(dom, wl), "algorithmic_bytes": pb[dom],
                "launch_ms": passes[dom]}

    forms_out = {}
    for form, rr in res.items():
        prec = 0 if form == "fp64" else 1
        sp = a.steps / rr["wall_s"]
        forms_out[form] = {
            "steps_per_s": sp,
            "ms_per_step": 1e3 * rr["wall_s"] / a.steps,
            "grid_points_steps_per_s": sp * M ** 3,
            "hbm_pct_model": 100.0 * model_bytes(N, prec, les=a.les) * sp / PEAK_HBM,
            "bytes_per_step_algorithmic": sum(pass_bytes(N, prec, les=a.les, fused_epilogue=fused(rr)).values()),
            "fwd_y_fused_with_epilogue": fused(rr) >= 1,
            "next_inv_y_fused": fused(rr) == 2,
            "pass_ms": rr["passes"],
            "ek_total": rr["scalars"]["ek"],
            "epsilon": rr["scalars"]["nu"] * rr["scalars"]["om"],
            "skewness": rr["scalars"]["skewness"],
        }
    split_err = None
    if "fp64" in res and "fp32" in res:
        import numpy as np
        e64, e32 = res["fp64"]["ek"], res["fp32"]["ek"]
        s64, s32 = res["fp64"]["scalars"], res["fp32"]["scalars"]
        split_err = {
            "steps": a.warmup + a.steps,
            "ek_rel_l2": float(np.linalg.norm(e32 - e64) / np.linalg.norm(e64)),
            "epsilon_rel": abs(s32["nu"] * s32["om"] - s64["nu"] * s64["om"]) / (s64["nu"] * s64["om"]),
            "skewness_abs": abs(s32["skewness"] - s64["skewness"]),
            "reference": "FP64 plain GPU form (which matches the FP64 oracle to 1e-12 at test sizes)",
        }

    ens = None
    if a.ensemble_requests > 0:
        print("bench: ensemble mix, %d requests" % a.ensemble_requests, file=sys.stderr, flush=True)
        ens = run_ensemble(a.ensemble_requests, a.seed, rate=a.ensemble_rate, chunk=a.chunk)
        print("bench: ensemble %.1f s wall, p50 %.2f s, p99 %.2f s" % (ens["wall_s"], ens["latency_p50_s"],
              ens["latency_p99_s"]), file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cN = a.cpu_N or N
        print("bench: cpu baseline, oracle %d step(s) at N=%d" % (a.cpu_steps, cN), file=sys.stderr, flush=True)
        dt, ts, used = cpu_baseline(cN, threads, a.cpu_steps, les=a.les)
        Mc = dims(cN)[0]
        scale = (M ** 3 * math.log2(M ** 3)) / (Mc ** 3 * math.log2(Mc ** 3))
        what = "%d step(s) of the same %s setup at N=%d (FFT grid %d^3), %s %.2f s/step (%s), 1 process x %d OpenMP " \
               "threads on %s, field pages touched before timing" % (
                   a.cpu_steps, "LES" if a.les else "forced-DNS", cN, Mc, "median" if a.cpu_steps > 1 else "measured",
                   dt, ", ".join("%.2f" % t for t in ts), used, cpu_model())
        if cN != N:
            what += "; value scaled to N=%d by the P*log2(P) ratio %.2f (an extrapolation)" % (N, scale)
        cpu = {"value": 1.0 / (dt * scale), "unit": "steps/s", "cores": used, "kind": "port", "sample": what,
               "measured_steps_per_s_at_sample": 1.0 / dt}

    line = {
        "metric": METRIC,
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 split" if main_form == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": ("Smagorinsky LES %d^3 retained modes (FFT grid %d^3), nu=%g, C_s=%g, Delta=2pi/%d, "
                                "CFL C_At=0.5, k0=%g u_rms=1 seed %d" % (N, M, LES_NU, LES_CS, N, LES_K0, a.seed))
                   if a.les else
                   ("forced DNS %d^3 retained modes (FFT grid %d^3), nu=%g, constant power 0.1 "
                    "on 1<=K<=2, CFL C_At=0.5, k0=4 u_rms=1 seed %d" % (N, M, workload_nu(N), a.seed)),
                   "N": N, "M": M, "form": main_form, "parallelism": "1 GPU"},
        "forms": forms_out,
        "fp32_split_speedup": (forms_out["fp32"]["steps_per_s"] / forms_out["fp64"]["steps_per_s"]
                               if "fp32" in forms_out and "fp64" in forms_out else None),
        "fp32_split_error": split_err,
        "roofline": roof,
        "copy_peak": copy,
        "cpu_baseline": cpu,
        "ensemble": ens,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
