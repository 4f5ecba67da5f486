"""Replays the reference's QA runs (src/MAIN.cpp:127-442) on any backend that
offers the HIT-like interface used below (oracle.Oracle or the product's
spns_amd.Solver), and formats the QA log rows as src/MAIN.cpp:440 does.

Parameter derivation follows src/MAIN.cpp:130-357; the loop follows
src/MAIN.cpp:413-442.
"""
import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATA_FILES = {
    "DATA/vel_32x32x32.dat": "vel_32x32x32.dat",
    "DATA/vel_48x48x48.bin": "vel_48x48x48.bin",
}


def load_fixtures():
    with open(os.path.join(GOLDEN, "qa_ref.json")) as f:
        return json.load(f)


def derive(params):
    """Key/values of a parameters file -> HIT constructor arguments + driver flags."""
    kv = {k.lower(): v for k, v in params.items()}
    g = lambda k, d=None: kv.get(k.lower(), d)
    d = {}
    Nx, Ny, Nz = int(g("Nx")), int(g("Ny")), int(g("Nz"))
    d["isReLambda"] = g("Reynolds_lambda") is not None
    d["ReLambda"] = float(g("Reynolds_lambda", 0.0))
    d["ReLambda_Freq"] = int(g("Reynolds_lambda_Freq", 0))
    d["isReLambda_iter"] = d["isReLambda"] and d["ReLambda_Freq"] > 0
    d["nu"] = float(g("nu", 1.0))
    d["Final_Time"] = float(g("Final_Time"))
    Mx, My, Mz = (3 * Nx) // 2, (3 * Ny) // 2, (3 * Nz) // 2
    fileset = all(g(k) is not None for k in ("Nx_file", "Ny_file", "Nz_file"))
    d["binary"] = g("Binary_Input_Filename") if fileset else None
    d["ascii"] = g("ASCII_Input_Filename") if fileset and d["binary"] is None else None
    d["Nfile"] = int(g("Nx_file", 0))
    d["isForcedEk"] = all(g(k) is not None for k in ("InputEkFilename", "Last_K", "Final_Forced_Time"))
    if d["isForcedEk"]:
        d["EkFile"] = g("InputEkFilename")
        d["Last_K"] = int(g("Last_K"))
        fft = float(g("Final_Forced_Time"))
        d["Final_Forced_Time"] = fft if fft > 0 else d["Final_Time"]
        d["nullify"] = int(g("Nullify_MissingK", 0)) != 0
    Lx, Ly, Lz = int(g("Lx", 1)), int(g("Ly", 1)), int(g("Lz", 1))
    notcubic = False
    if Lx > 1 or Ly > 1 or Lz > 1:
        lc = math.lcm(Lx, math.lcm(Ly, Lz))
        Lx, Ly, Lz = lc // Lx, lc // Ly, lc // Lz
        if Lx > 1 or Ly > 1 or Lz > 1:
            notcubic = True
            Mx = Mx + 1 if (Lx > 1 and Mx % 2) else Mx
            My = My + 1 if (Ly > 1 and My % 2) else My
            Mz = Mz + 1 if (Lz > 1 and Mz % 2) else Mz
            f = lambda n, L: n * L if n % 2 == 0 else (n - 1) * L + 1
            Nx, Ny, Nz = f(Nx, Lx), f(Ny, Ly), f(Nz, Lz)
            Mx, My, Mz = f(Mx, Lx), f(My, Ly), f(Mz, Lz)
    else:
        Lx = Ly = Lz = 1
    d["ctor"] = dict(Nx=Nx, Ny=Ny, Nz=Nz, Mx=Mx, My=My, Mz=Mz, nu=d["nu"],
                     C_Smag=float(g("C_Smag", 0.0)), C_At=float(g("C_At", 1.0)),
                     Lx=Lx, Ly=Ly, Lz=Lz,
                     omega=tuple(float(g(k, 0.0)) for k in ("omega_x", "omega_y", "omega_z")),
                     isNotCubic=notcubic,
                     isSelfAdaptive=int(g("SelfAdaptiveTimestep", 0)) == 1,
                     isCCC=int(g("ComplexConjugateCorrection", 0)) == 1)
    return d


def read_field(d):
    n = d["Nfile"]
    if d["ascii"]:
        a = np.loadtxt(os.path.join(GOLDEN, DATA_FILES[d["ascii"]]))
        return [a[:, c].reshape(n, n, n) for c in range(3)]
    if d["binary"]:
        a = np.fromfile(os.path.join(GOLDEN, DATA_FILES[d["binary"]]), dtype=np.float64)
        P = n ** 3
        return [a[c * P:(c + 1) * P].reshape(n, n, n) for c in range(3)]
    return None


def read_ek(d):
    path = os.path.join(GOLDEN, "Forced_Energy_Cascade_K%d.dat" % d["Last_K"])
    ek = np.zeros(d["Last_K"] + 1)
    for line in open(path):
        t = line.split()
        if len(t) >= 2:
            ek[int(t[0])] = float(t[1])
    return ek


def fmt_row(s, it):
    return ["%f" % s["time"], "%d" % it, "%e" % s["dt"], "%e" % s["ek"], "%e" % s["om"], "%.3f" % s["relambda"]]


def run(params, make_solver, max_steps=None):
    """make_solver(**ctor) -> backend.  Returns the list of formatted rows."""
    d = derive(params)
    s = make_solver(**d["ctor"])
    field = read_field(d)
    ekf = ensf = None
    if field is not None:
        ekf, ensf = s.init_from_physical(*field)
    else:
        s.init_k41(4)
    s.finish_init(d["isReLambda"])
    fek = read_ek(d) if d["isForcedEk"] else None
    if d["isReLambda"]:
        if d["isForcedEk"]:
            et = ot = 0.0
            for k in range(1, d["Last_K"] + 1):
                et += fek[k]
                ot += 2 * k * k * fek[k]
            nu = et / d["ReLambda"] * math.sqrt(20.0 / 3.0 / ot)
        elif ekf is not None:
            nu = ekf / d["ReLambda"] * math.sqrt(20.0 / 3.0 / ensf)
        else:
            nu = 1.0e-3
        s.set_nu(nu)
        s.recalc_timestep()
    rows = []
    it = 0
    while s.scalars()["time"] < d["Final_Time"]:
        s.step()
        if it == 0 and d["isReLambda"] and not d["isForcedEk"] and field is None:
            s.forcing_relambda(d["ReLambda"])
            s.recalc_timestep()
        if d["isForcedEk"] and s.scalars()["time"] < d["Final_Forced_Time"]:
            s.forcing_energy_cascade(fek, d["Last_K"], d["nullify"])
        if d["isReLambda_iter"] and it % d["ReLambda_Freq"] == 0:
            s.forcing_relambda(d["ReLambda"])
            s.recalc_timestep()
        it += 1
        s.relambda()
        rows.append(fmt_row(s.scalars(), it))
        if max_steps and it >= max_steps:
            break
    return rows, s
