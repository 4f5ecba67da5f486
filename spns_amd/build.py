"""Builds the HIP library spns_amd/libspns.so (gfx950) and the driver
spns_amd/spns_run in-tree with hipcc.

    python -m spns_amd.build [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspns.so")
RUN = os.path.join(HERE, "spns_run")
ARCH = os.environ.get("SPNS_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions stay off: the compiler must not form v_pk_* FP32 ops
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

UNITS = [
    ("spns_kernels.hip", "k1.o", ["-DSPNS_PART=1"]),
    ("spns_kernels.hip", "k2.o", ["-DSPNS_PART=2"]),
    ("spns_kernels.hip", "k3.o", ["-DSPNS_PART=3"]),
    ("spns_kernels.hip", "k4.o", ["-DSPNS_PART=4"]),
    ("spns_ctx.hip", "ctx.o", []),
    ("hit.cpp", "hit.o", []),
    ("params.cpp", "params.o", []),
]


def _compile(src, obj, extra):
    cmd = [HIPCC, "--offload-arch=" + ARCH] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), msg))
    return msg


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u[0]))]
    jobs = jobs or min(len(units), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_compile, *u) for u in units]
        for f in futs:
            m = f.result()
            if verbose and m:
                print(m)
    objs = [os.path.join(BUILD, u[1]) for u in units]
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    drv = os.path.join(CSRC, "spns_run.cpp")
    if os.path.exists(drv):
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-o", RUN, drv, "-L" + HERE, "-lspns",
                        "-Wl,-rpath,$ORIGIN"], check=True)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
