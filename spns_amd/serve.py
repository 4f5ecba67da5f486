"""Serving loop: continuous batching of independent HIT runs on one GPU.

Requests (independent runs of mixed grids, viscosities, Smagorinsky
constants and rotation rates) arrive over time.  The loop
  * admits a request when it has arrived and its device footprint
    (spns_device_bytes) fits the HBM budget left by the runs in flight (the
    budget from spns_device_memory);
  * keeps every admitted run resident: state, history and work buffers stay
    in HBM from admission to completion;
  * batched (the default): seats each run in a slot of its shape class's
    Batch (grid, precision, DNS or LES) and advances every batch in chunks of
    hipGraph-captured steps, each pass one kernel launch over all the
    batch's runs; batches of different classes run concurrently on their own
    streams, and runs join and leave batches between chunks;
  * reads each run's status after every chunk from pinned copies the batch
    makes on its own stream, and a finished run's results the same way, so
    the loop never waits on a device read behind other streams' work;
  * retires runs that reached their step count (E(k), energy, dissipation,
    skewness, state) and keeps their contexts for later runs of the same
    class (spns_reconfigure) instead of freeing device memory, which would
    wait for all work on the device;
  * retires a run that diverged (NaN/Inf: SPNS_ENONFINITE) or failed to
    start, with an error result, and keeps the others going: one bad request
    never takes the loop down (the reference's crash() aborts every rank
    instead, src/util.cpp:93-116).
Per-run mode (batched=False): one graph per run on its own stream, polled
with spns_query, shortest remaining work first with aging, and a timing
autotuner (Tuner) for the z-pass kernel of each shape class.
Latency = completion time - arrival time; the loop reports p50/p99 beside
throughput.

Runs never share state: a run's result is the same alone and in the mix, bit
for bit, batched or not (tests/test_gpu_serve.py, tests/test_gpu_batch.py).
"""
import concurrent.futures as cf
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import PREC_FP32_SPLIT, Batch, Pinned, Solver, SpnsError, device_bytes, device_memory

GRIDS = [(32, 32, 32), (64, 64, 64), (96, 64, 64), (128, 128, 128), (192, 128, 128)]


@dataclass
class Request:
    rid: int
    seed: int
    grid: tuple
    C_Smag: float
    nu: float
    omega_z: float
    steps: int
    arrival: float
    k0: float = 4.0
    fixed_dt: float = 0.0  # > 0: a fixed time step instead of the CFL step
    # filled by the server
    admitted: float = math.nan
    finished: float = math.nan
    result: dict = field(default_factory=dict)
    error: str = ""  # set when the run failed (diverged, or could not start)

    def solver_args(self, precision):
        Nx, Ny, Nz = self.grid
        return dict(Nx=Nx, Ny=Ny, Nz=Nz, nu=self.nu, C_Smag=self.C_Smag, C_At=0.5,
                    omega=(0.0, 0.0, self.omega_z), precision=precision, fixed_dt=self.fixed_dt)


def draw_requests(n, rate=20.0, seed=0, steps=(100, 2000)):
    """The ensemble mix: Poisson arrivals at `rate`/s; each request drawn from
    its own seed: grid uniform over GRIDS, DNS or Smagorinsky (C_s uniform in
    [0.1, 0.2]), nu log-uniform in [1e-4, 1e-2], Omega_z in {0, 1, 5},
    steps uniform in [steps[0], steps[1]]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        r = np.random.default_rng([seed, i])
        grid = GRIDS[int(r.integers(len(GRIDS)))]
        les = bool(r.integers(2))
        out.append(Request(rid=i, seed=int(r.integers(1 << 31)), grid=grid,
                           C_Smag=float(r.uniform(0.1, 0.2)) if les else 0.0,
                           nu=float(10 ** r.uniform(-4, -2)), omega_z=float([0.0, 1.0, 5.0][int(r.integers(3))]),
                           steps=int(r.integers(steps[0], steps[1] + 1)), arrival=t))
    return out


class Tuner:
    """Picks the z-pass kernel variant of a shape class (0 staged inputs,
    1 block per pencil, 2 wave per pencil for DNS) by timing a few steps of
    each on a throwaway run."""

    def __init__(self, steps=4):
        self.steps = steps
        self.choice = {}

    @staticmethod
    def key(args):
        return (args["Nx"], args["Ny"], args["Nz"], args.get("precision", 0), args.get("C_Smag", 0.0) > 0)

    def prepare(self, arg_list):
        """Tunes every shape class of arg_list (call while the device is
        idle, before any run is admitted: the timings then see no other
        work, and the throwaway runs' memory is free again afterwards)."""
        for args in arg_list:
            self.variant(args)

    def variant(self, args):
        key = self.key(args)
        if key in self.choice:
            return self.choice[key]
        best, best_ms = 0, math.inf
        for v in ((0, 1) if args.get("C_Smag", 0.0) > 0 else (0, 1, 2)):
            s = Solver(zpass_variant=v, **args)
            s.init_random(seed=0)
            s.finish_init(False)
            s.step(1)
            s.synchronize()
            s.step(self.steps)
            ms = s.last_step_ms()
            s.close()
            if ms < best_ms:
                best, best_ms = v, ms
        self.choice[key] = best
        return best


def step_cost(req):
    """Relative cost of one step of a request: its FFT grid points (x 1.4 for
    LES, nine forward fields instead of three), at least 1e6 (small grids are
    bound by launch latency, not work)."""
    Nx, Ny, Nz = req.grid
    pts = (3 * Nx // 2) * (3 * Ny // 2) * (3 * Nz // 2)
    return max(pts * (1.4 if req.C_Smag > 0 else 1.0), 1e6)


def eligible_runs(active, inflight_work, min_inflight, max_wait=None):
    """Indices into active (dicts with "req", "done" and "waited") of the runs
    that may get their next chunk: all of them when inflight_work is None,
    else the runs in ascending order of remaining work (step_cost x steps
    left) while their per-step costs fit inflight_work, and at least
    min_inflight.  Aging: a run that has been passed over max_wait times in
    a row ("waited", counted by the caller) is chosen as well, whatever its
    remaining work, on top of that budget, so a long run cannot starve
    behind a stream of short ones."""
    if inflight_work is None:
        return set(range(len(active)))
    def overdue(i):
        return max_wait is not None and active[i].get("waited", 0) >= max_wait
    order = sorted(range(len(active)),
                   key=lambda i: (not overdue(i),
                                  step_cost(active[i]["req"]) * (active[i]["req"].steps - active[i]["done"]), i))
    chosen, budget, n = set(), 0.0, 0
    for i in order:
        if overdue(i):  # on top of the budget and of min_inflight
            chosen.add(i)
            continue
        c = step_cost(active[i]["req"])
        if n >= min_inflight and budget + c > inflight_work:
            break
        chosen.add(i)
        budget += c
        n += 1
    return chosen


def fft_points(req):
    Nx, Ny, Nz = req.grid
    return (3 * Nx // 2) * (3 * Ny // 2) * (3 * Nz // 2)


def class_key(req, precision):
    """Shape class of a request: runs of one class share batched launches."""
    return (tuple(req.grid), precision, req.C_Smag > 0)


class Server:
    """batched=True (the default): runs of one shape class (grid, precision,
    DNS or LES) share a Batch of nslots slots, and each pass of a step is one
    kernel launch over every run of the batch; a chunk of up to `chunk`
    steps is one captured graph, and the batches of different classes run
    concurrently on their own streams.  Within a batch every run advances
    min(chunk, steps left) per chunk; a class whose batches are full opens
    another batch.  Runs wait (admitted, initialised) for a slot until their
    batch is between chunks.

    batched=False: one graph per run on its own stream.  inflight_work: None
    advances every admitted run as soon as its previous chunk finishes
    (processor sharing).  A number (the default, 2e7) gives the next chunks
    to the runs with the least remaining work (step_cost x steps left) first,
    as many as fit that per-step cost budget (at least min_inflight), so
    short runs finish early and long ones use the slack."""

    def __init__(self, hbm_budget=None, precision=PREC_FP32_SPLIT, chunk=20, max_active=128, tuner=None,
                 inflight_work=2e7, min_inflight=6, max_wait=16, device=0, batched=True, nslots=16,
                 keep_contexts=True, admit_per_pass=1):
        if hbm_budget is None:
            hbm_budget = device_memory(device)[0] * 0.8
        self.budget = hbm_budget
        self.precision = precision
        self.chunk = chunk
        self.max_active = max_active
        self.tuner = tuner
        self.inflight_work = inflight_work
        self.min_inflight = max(1, int(min_inflight))
        self.max_wait = max_wait
        self.device = device
        self.batched = batched
        self.nslots = nslots
        self.keep_contexts = keep_contexts
        self.admit_per_pass = max(1, int(admit_per_pass))
        self.used = 0.0
        self.reused = 0
        self._kept = {}
        # a batch's first chunk captures its graph; an allocation on another
        # thread meanwhile invalidates the capture, so the two exclude each other
        self._capture_lock = threading.Lock()

    # Finished runs' contexts are kept per shape class and reused by later
    # runs of that class (spns_reconfigure): freeing device memory waits for
    # all work on the device, which would stall the host, and every batch
    # with it, at each retirement.  Kept contexts count against the HBM
    # budget and are freed only when a new run needs their memory.
    def _pool_key(self, args, v):
        return (args["Nx"], args["Ny"], args["Nz"], args["precision"], args["C_Smag"] > 0.0, v)

    @staticmethod
    def _close_entry(e):
        e["solver"].close()
        for p in e.get("pin") or ():
            p.close()

    def _evict(self, need):
        """Frees kept contexts (oldest first) until need bytes fit the budget."""
        while self.used + need > self.budget and self._kept:
            key = next(iter(self._kept))
            e = self._kept[key].pop(0)
            if not self._kept[key]:
                del self._kept[key]
            self._close_entry(e)
            self.used -= e["bytes"]

    def _claim(self, req, now):
        """The bookkeeping half of admitting a run (the loop's thread): a kept
        context of the run's class, or the bytes of a new one, on the budget."""
        args = req.solver_args(self.precision)
        # batched launches use the default (staged) z-pass kernels
        v = self.tuner.variant(args) if (self.tuner and not self.batched) else 0
        key = self._pool_key(args, v)
        nbytes = device_bytes(zpass_variant=v, **args)
        kept = self._kept.get(key)
        e = None
        if kept:
            e = kept.pop(0)
            if not kept:
                del self._kept[key]
            nbytes = e["bytes"]
            self.reused += 1
        else:
            self.used += nbytes
        req.admitted = now
        return {"req": req, "args": args, "v": v, "key": key, "bytes": nbytes, "entry": e}

    def _init(self, c):
        """The device half (any thread; it touches only this run's context):
        create or reconfigure the context, seed the field, first antitransform.
        On an error the context is closed and the error returned."""
        req, e, pin, s = c["req"], c["entry"], None, None
        with self._capture_lock:  # no allocation while a batch captures its graph
            return self._init_locked(c, req, e, pin, s)

    def _init_locked(self, c, req, e, pin, s):
        try:
            if e is not None:
                s, pin = e["solver"], e.get("pin")
                s.reconfigure(**c["args"])
            else:
                s = Solver(zpass_variant=c["v"], **c["args"])
            s.init_random(seed=req.seed, k0=req.k0)
            s.finish_init(False)
        except SpnsError as err:
            if s is not None:
                self._close_entry({"solver": s, "pin": pin})
            return err
        return {"req": req, "solver": s, "done": 0, "waited": 0, "bytes": c["bytes"], "key": c["key"], "pin": pin}

    def _start(self, req, now):
        c = self._claim(req, now)
        a = self._init(c)
        if isinstance(a, SpnsError):
            self.used -= c["bytes"]
            raise a
        return a

    def _release(self, a, keep):
        s = a.get("solver")
        if s is None:
            return
        e = {"solver": s, "bytes": a["bytes"], "pin": a.get("pin")}
        if keep and self.keep_contexts:
            self._kept.setdefault(a["key"], []).append(e)
        else:
            self._close_entry(e)
            self.used -= a["bytes"]

    def _finish_from(self, a, sc, state, ek, now):
        """Result of a batched run from its fetched (pinned) copies."""
        req = a["req"]
        req.result = {"ek": np.array(ek), "energy": sc["ek"], "epsilon": sc["nu"] * sc["om"],
                      "skewness": sc["skewness"], "time": sc["time"], "steps": sc["iter"], "state": np.array(state)}
        req.finished = now
        self._release(a, keep=True)

    def _finish(self, a, now):
        s, req = a["solver"], a["req"]
        try:
            sc = s.scalars()
            req.result = {"ek": s.spectrum(), "energy": sc["ek"], "epsilon": sc["nu"] * sc["om"],
                          "skewness": sc["skewness"], "time": sc["time"], "steps": sc["iter"], "state": s.get_state()}
        except SpnsError as e:
            self._fail(a, e, now)
            return
        req.finished = now
        self._release(a, keep=True)

    def _fail(self, a, err, now):
        req = a["req"]
        req.error = str(err)
        req.result = {"error_code": getattr(err, "code", -1), "steps": a.get("done", 0)}
        req.finished = now
        self._release(a, keep=False)  # a diverged run's buffers are not reused

    def _close_kept(self):
        for lst in self._kept.values():
            for e in lst:
                self._close_entry(e)
                self.used -= e["bytes"]
        self._kept = {}

    @staticmethod
    def _stats(done, wall, steps_done):
        ok = [r for r in done if not r.error]
        lat = np.array([r.finished - r.arrival for r in ok])
        return {
            "requests": len(done),
            "failed": len(done) - len(ok),
            "wall_s": wall,
            "runs_per_s": len(ok) / wall,
            "steps_per_s": steps_done / wall,
            "grid_point_steps_per_s": sum(fft_points(r) * r.steps for r in ok) / wall,
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else math.nan,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else math.nan,
            "latency_mean_s": float(lat.mean()) if len(lat) else math.nan,
        }

    def run(self, requests, keep_state=False):
        """Serves the requests (arrival times relative to the call); returns
        (stats, the requests in completion order)."""
        self.reused = 0
        stats, done = (self._run_batched if self.batched else self._run_unbatched)(requests, keep_state)
        stats["contexts_reused"] = self.reused
        return stats, done

    def _run_batched(self, requests, keep_state):
        pending = sorted(requests, key=lambda r: r.arrival)
        classes = {}  # class key -> list of groups {"batch", "slots": {slot: run}, "waiting": [run], "busy"}
        done = []
        n_runs, steps_done = 0, 0
        host = {"start": 0.0, "retire": 0.0, "attach": 0.0, "launch": 0.0, "poll": 0.0, "idle": 0.0}
        clock = time.perf_counter
        t0 = time.perf_counter()
        # initialisations (context creation, seeded field, first antitransform:
        # tens of ms each) run on a worker thread, so the loop keeps relaunching
        # the batches meanwhile; each touches only its own context and stream
        worker = cf.ThreadPoolExecutor(1)
        inits = []  # (future, claim)

        def retire(g, slot, a, failed_err=None):
            nonlocal n_runs
            g["batch"].detach(slot)
            del g["slots"][slot]
            now = time.perf_counter() - t0
            if failed_err is not None:
                self._fail(a, failed_err, now)
            else:
                st, ek = a["pin"]
                self._finish_from(a, a["final"], st.array, ek.array, now)
                if not keep_state:
                    a["req"].result.pop("state", None)
            n_runs -= 1
            done.append(a["req"])

        try:
            while pending or n_runs:
                now = time.perf_counter() - t0
                progressed = False
                # admission by HBM footprint, in arrival order, at most
                # admit_per_pass runs per pass (an initialisation takes tens of
                # ms; the running batches are relaunched in between); the run
                # is initialised now and waits for a slot of its class's batch
                admitted = 0
                while (pending and pending[0].arrival <= now and n_runs < self.max_active
                       and admitted < self.admit_per_pass):
                    admitted += 1
                    r = pending[0]
                    args = r.solver_args(self.precision)
                    need = 0.0 if self._kept.get(self._pool_key(args, 0)) else device_bytes(**args)
                    self._evict(need)
                    if self.used + need > self.budget and n_runs:
                        break
                    pending.pop(0)
                    tc = clock()
                    c = self._claim(r, now)
                    inits.append((worker.submit(self._init, c), c))
                    host["start"] += clock() - tc
                    n_runs += 1
                    progressed = True
                # initialised runs wait for a slot of their class's batch
                for item in [x for x in inits if x[0].done()]:
                    inits.remove(item)
                    fut, c = item
                    r = c["req"]
                    a = fut.result()
                    if isinstance(a, SpnsError):
                        self.used -= c["bytes"]
                        self._fail({"req": r}, a, time.perf_counter() - t0)
                        done.append(r)
                        n_runs -= 1
                        continue
                    groups = classes.setdefault(class_key(r, self.precision), [])
                    g = next((g for g in groups if len(g["slots"]) + len(g["waiting"]) < self.nslots), None)
                    if g is None:
                        g = {"batch": Batch(self.nslots, self.chunk, self.device), "slots": {}, "waiting": [], "busy": False}
                        groups.append(g)
                    g["waiting"].append(a)
                    progressed = True
                # every batch between chunks: retire finished or diverged
                # runs, seat waiting ones, launch the next chunk
                for groups in classes.values():
                    for g in groups:
                        if g["busy"]:
                            tc = clock()
                            busy = not g["batch"].query()
                            host["poll"] += clock() - tc
                            if busy:
                                continue
                            g["busy"] = False
                        tc = clock()
                        for slot, a in list(g["slots"].items()):
                            if "final" in a:  # its fetched results have landed
                                retire(g, slot, a)
                                progressed = True
                                continue
                            try:
                                sc = g["batch"].slot_scalars(slot)  # SPNS_ENONFINITE if it diverged
                            except SpnsError as e:
                                retire(g, slot, a, e)
                                progressed = True
                                continue
                            if a["done"] >= a["req"].steps:
                                # queue the copies of its results on the batch's stream;
                                # it retires when they have landed (after the next chunk)
                                s = a["solver"]
                                if a.get("pin") is None:
                                    a["pin"] = (Pinned((3, s.nmodes), np.complex128), Pinned((s.nbins,)))
                                g["batch"].fetch(slot, a["pin"][0].array, a["pin"][1].array)
                                a["final"] = sc
                                progressed = True
                        host["retire"] += clock() - tc
                        tc = clock()
                        while g["waiting"]:
                            free = g["batch"].free_slots()
                            if not free:
                                break
                            a = g["waiting"].pop(0)
                            g["batch"].attach(free[0], a["solver"])
                            g["slots"][free[0]] = a
                        host["attach"] += clock() - tc
                        plan = [0] * self.nslots
                        for slot, a in g["slots"].items():
                            if "final" not in a:
                                plan[slot] = min(self.chunk, a["req"].steps - a["done"])
                        if any(plan):
                            tc = clock()
                            if g.get("captured"):
                                g["batch"].step(plan)
                            else:
                                with self._capture_lock:
                                    g["batch"].step(plan)
                                g["captured"] = True
                            host["launch"] += clock() - tc
                            g["busy"] = True
                            for slot, a in g["slots"].items():
                                a["done"] += plan[slot]
                                steps_done += plan[slot]
                            progressed = True
                        elif any("final" in a for a in g["slots"].values()):
                            g["busy"] = True  # only fetches in flight
                if not progressed:
                    tc = clock()
                    if not n_runs and pending:
                        time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                    else:
                        time.sleep(2e-4)
                    host["idle"] += clock() - tc
        finally:
            worker.shutdown(wait=True)
            for fut, c in inits:  # only after an exception in the loop
                a = fut.result()
                if not isinstance(a, SpnsError):
                    self._release(a, keep=False)
            for groups in classes.values():
                for g in groups:
                    g["batch"].close()
        wall = time.perf_counter() - t0
        self._close_kept()
        stats = self._stats(done, wall, steps_done)
        stats["host_s"] = {k: round(v, 3) for k, v in host.items()}
        return stats, done

    def _run_unbatched(self, requests, keep_state):
        pending = sorted(requests, key=lambda r: r.arrival)
        if self.tuner:
            self.tuner.prepare([r.solver_args(self.precision) for r in pending])
        active, done = [], []
        t0 = time.perf_counter()
        steps_done = 0
        while pending or active:
            now = time.perf_counter() - t0
            # admission by HBM footprint, in arrival order
            while pending and pending[0].arrival <= now and len(active) < self.max_active:
                r = pending[0]
                args = r.solver_args(self.precision)
                v = self.tuner.variant(args) if self.tuner else 0
                need = 0.0 if self._kept.get(self._pool_key(args, v)) else device_bytes(zpass_variant=v, **args)
                self._evict(need)
                if self.used + need > self.budget and active:
                    break
                pending.pop(0)
                try:
                    a = self._start(r, now)
                except SpnsError as e:
                    self._fail({"req": r}, e, now)
                    done.append(r)
                    continue
                active.append(a)
            if not active:
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                continue
            # every run whose last chunk has finished gets its next chunk (or
            # retires); the others keep running: no run waits for another
            eligible = eligible_runs(active, self.inflight_work, self.min_inflight, self.max_wait)
            progressed = False
            still = []
            for ia, a in enumerate(active):
                s = a["solver"]
                try:
                    if a["done"] > 0 and not s.query():
                        still.append(a)
                        continue
                except SpnsError as e:  # diverged (or a device error) in its last chunk
                    self._fail(a, e, time.perf_counter() - t0)
                    done.append(a["req"])
                    progressed = True
                    continue
                if a["done"] >= a["req"].steps:
                    self._finish(a, time.perf_counter() - t0)
                    if not keep_state:
                        a["req"].result.pop("state", None)
                    done.append(a["req"])
                    progressed = True
                    continue
                if ia not in eligible:
                    a["waited"] += 1
                    still.append(a)  # waits for the shorter runs
                    continue
                n = min(self.chunk, a["req"].steps - a["done"])
                s.step(n)
                a["done"] += n
                a["waited"] = 0
                steps_done += n
                progressed = True
                still.append(a)
            active = still
            if not progressed:
                time.sleep(2e-4)
        wall = time.perf_counter() - t0
        self._close_kept()
        return self._stats(done, wall, steps_done), done


def run_alone(req, precision=PREC_FP32_SPLIT, zpass_variant=0):
    """Runs one request by itself (reference for the alone-vs-mix check)."""
    s = Solver(zpass_variant=zpass_variant, **req.solver_args(precision))
    s.init_random(seed=req.seed, k0=req.k0)
    s.finish_init(False)
    s.step(req.steps)
    st = s.get_state()
    sc = s.scalars()
    ek = s.spectrum()
    s.close()
    return {"state": st, "ek": ek, "energy": sc["ek"], "time": sc["time"]}
