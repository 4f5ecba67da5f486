"""spns_run, the MAIN.cpp-equivalent driver: parameter errors on the CPU
(src/MAIN.cpp:98, 174-175, src/parser.cpp:364-479) and, on the GPU, the
reference's QA logs reproduced from its own parameters files."""
import os
import re
import shutil
import subprocess

import pytest

import qa_driver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "spns_amd", "spns_run")
ROW = re.compile(r"Time: ([\d.]+),\s+iter:\s+(\d+),\s+Dt: (\S+),\s+Ek: (\S+),\s+Om: (\S+),\s+ReLambda:\s+(\S+)")

needs_run = pytest.mark.skipif(not os.path.exists(RUN), reason="spns_run not built")


def write_params(d, params, extra=None):
    lines = ["%s\t%s" % (k, v) for k, v in params.items()]
    for k, v in (extra or {}).items():
        lines.append("%s\t%s" % (k, v))
    path = os.path.join(d, "params")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


@needs_run
@pytest.mark.parametrize("text,msg", [
    ("Ny 32\nNz 32\nFinal_Time 1\nnu 0.01\n", "Nx is missing"),
    ("Nx 32\nNy 32\nNz 32\nFinal_Time 1\n", "Either kinematic viscosity or Reynolds lambda"),
    ("Nx 32\nNy 32\nNz 32\nFinal_Time 1\nnu 0.01\nReynolds_lambda 100\n", "cannot be defined simoultaniously"),
    ("Nx 32\nnx 16\nNy 32\nNz 32\nFinal_Time 1\nnu 0.01\n", "double definition"),
])
def test_parameter_errors(tmp_path, text, msg):
    p = tmp_path / "params"
    p.write_text(text)
    r = subprocess.run([RUN, str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert msg in r.stderr


@needs_run
def test_missing_parameters_file(tmp_path):
    r = subprocess.run([RUN, str(tmp_path / "nope")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Parameters file is missing" in r.stderr


FIX = qa_driver.load_fixtures()


@pytest.mark.gpu
@needs_run
@pytest.mark.parametrize("name", sorted(FIX))
def test_driver_reproduces_qa_log(tmp_path, name):
    fx = FIX[name]
    params = dict(fx["params"])
    os.makedirs(tmp_path / "DATA", exist_ok=True)
    for src, dst in qa_driver.DATA_FILES.items():
        shutil.copy(os.path.join(qa_driver.GOLDEN, dst), tmp_path / src)
    if "InputEkFilename" in params:
        shutil.copy(os.path.join(qa_driver.GOLDEN, "Forced_Energy_Cascade_K%s.dat" % params["Last_K"]),
                    tmp_path / "DATA" / "Forced_Energy_Cascade.dat")
    extra = {"K41_Ranks": 4}  # the REF logs were made with 4 MPI ranks
    for k in ["Velocity_Physical", "Velocity_Fourier", "Energy_Cascade"]:
        params[k] = "0"
    path = write_params(str(tmp_path), params, extra)
    r = subprocess.run([RUN, path, "--qa"], capture_output=True, text=True, cwd=str(tmp_path), timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [list(m.groups()) for m in map(ROW.match, r.stdout.splitlines()) if m]
    assert rows == fx["rows"]
    finals = dict(l.split(":", 1) for l in r.stdout.splitlines() if l.startswith("Final "))
    assert {k.strip(): v.strip() for k, v in finals.items()} == fx["final"]


@pytest.mark.gpu
@needs_run
def test_driver_writes_invariant_spectra(tmp_path):
    """VelGradInv_Spectra / VelGradInv_Spectra_Freq (src/MAIN.cpp:311-317,
    460-469, 496-503): numbered QG QS RG RS V2 files every Freq-th iteration
    and the unnumbered ones at the end, "K<TAB>value" rows; the final files
    hold what the library returns for the same seeded run."""
    import numpy as np
    import spns_amd
    params = {"Nx": 16, "Ny": 16, "Nz": 16, "nu": 0.01, "C_At": 0.5, "Final_Time": 100.0, "Max_Iterations": 4,
              "Random_Seed": 3, "Random_k0": 4.0, "VelGradInv_Spectra": 1, "VelGradInv_Spectra_Freq": 2,
              "Spectra_Folder": str(tmp_path)}
    path = write_params(str(tmp_path), params)
    r = subprocess.run([RUN, path], capture_output=True, text=True, cwd=str(tmp_path), timeout=300)
    assert r.returncode == 0, r.stderr
    names = ("QG", "QS", "RG", "RS", "V2")
    for n in names:
        for i in (0, 1):
            assert (tmp_path / ("%s_%08d.dat" % (n, i))).exists()
        assert not (tmp_path / ("%s_%08d.dat" % (n, 2))).exists()
    s = spns_amd.Solver(Nx=16, nu=0.01, C_At=0.5, gradient_invariants=True)
    s.init_random(seed=3, k0=4.0)
    s.finish_init(False)
    s.step(4)
    ref = s.invariant_spectra()
    s.close()
    for f, n in enumerate(names):
        rows = [l.split("\t") for l in (tmp_path / ("%s.dat" % n)).read_text().splitlines()]
        assert [int(k) for k, _ in rows] == list(range(ref.shape[1]))
        got = np.array([float(v) for _, v in rows])
        assert np.allclose(got, ref[f], rtol=1e-5, atol=1e-6 * np.abs(ref[f]).max()), n


@pytest.mark.gpu
@needs_run
def test_driver_velocity_dumps_and_restart(tmp_path):
    """Velocity_Fourier / Velocity_Physical dumps (src/MAIN.cpp:471-484, 505-514,
    src/HIT_LES_output.cpp:20-115).  Fourier dumps at (iter+1) % Freq == 0 hold
    the current state; physical dumps at iter % Freq == 0 hold the field of the
    last step's start state, which is the state of the previous Fourier dump
    (the reference's u,v,w arrays between steps); the final physical dump
    antitransforms the current state first.  Each file is compared bit for bit
    with the library's own state and fields of the same seeded run, and the
    final physical dump read back as a binary input file restores the state."""
    import numpy as np
    import spns_amd
    params = {"Nx": 16, "Ny": 16, "Nz": 16, "nu": 0.01, "C_At": 0.5, "Final_Time": 100.0, "Max_Iterations": 4,
              "Random_Seed": 3, "Random_k0": 4.0, "Velocity_Fourier": 1, "Velocity_Four_Freq": 2,
              "Velocity_Physical": 1, "Velocity_Phys_Freq": 2, "Velocity_Four_Folder": str(tmp_path),
              "Velocity_Phys_Folder": str(tmp_path)}
    path = write_params(str(tmp_path), params)
    r = subprocess.run([RUN, path], capture_output=True, text=True, cwd=str(tmp_path), timeout=300)
    assert r.returncode == 0, r.stderr
    M = 24
    s = spns_amd.Solver(Nx=16, nu=0.01, C_At=0.5, keep_step_start=True)
    s.init_random(seed=3, k0=4.0)
    s.finish_init(False)
    states, fields = {}, {}
    for n in range(1, 5):
        s.step(1)
        states[n] = s.get_state().ravel()
        fields[n - 1] = s.velocity_physical_step_start()  # the field of state n-1
    final_phys = s.velocity_physical()
    s.close()
    four = lambda name: np.fromfile(tmp_path / name, dtype=np.complex128)
    phys = lambda name: np.fromfile(tmp_path / name, dtype=np.float64).reshape(3, M, M, M)
    assert np.array_equal(four("Velocity_Fourier_00000000.bin"), states[1])
    assert np.array_equal(four("Velocity_Fourier_00000001.bin"), states[3])
    assert np.array_equal(four("Velocity_Fourier.bin"), states[4])
    assert np.array_equal(phys("Velocity_00000000.bin"), fields[1])
    assert np.array_equal(phys("Velocity_00000001.bin"), fields[3])
    assert np.array_equal(phys("Velocity.bin"), final_phys)
    assert not (tmp_path / "Velocity_00000002.bin").exists()
    # restart from the final physical dump: the same state back
    g = spns_amd.Solver(Nx=16, nu=0.01, C_At=0.5)
    u = phys("Velocity.bin")
    g.init_from_physical(u[0], u[1], u[2])
    back = g.get_state().ravel()
    g.close()
    assert np.linalg.norm(back - states[4]) <= 1e-13 * np.linalg.norm(states[4])
