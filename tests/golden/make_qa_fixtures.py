"""Writes tests/golden/qa_ref.json from the reference's QA regression data.

Reads QA/TEST_*/params* and QA/TEST_*/REF/stdout.0 under the reference tree
(default /root/reference) and stores, per test, the parameter key/values and
the per-step log rows (Time, iter, Dt, Ek, Om, ReLambda) plus the Final lines.
The input files those tests use (vel_32x32x32.dat, vel_48x48x48.bin and the
two Forced_Energy_Cascade.dat spectra) are copied beside this script.
"""
import glob
import json
import os
import re
import shutil
import sys

ROW = re.compile(r"Time: ([\d.]+),\s+iter:\s+(\d+),\s+Dt: (\S+),\s+Ek: (\S+),\s+Om: (\S+),\s+ReLambda:\s+(\S+)")


def main(ref_root="/root/reference"):
    here = os.path.dirname(os.path.abspath(__file__))
    out = {}
    for tdir in sorted(glob.glob(os.path.join(ref_root, "QA", "TEST_*"))):
        name = os.path.basename(tdir)
        params = {}
        for line in open(glob.glob(os.path.join(tdir, "params*"))[0]):
            line = line.split("#")[0].strip()
            if line:
                t = line.split()
                params[t[0]] = t[1] if len(t) > 1 else ""
        rows, final = [], {}
        for line in open(os.path.join(tdir, "REF", "stdout.0"), errors="replace"):
            m = ROW.match(line)
            if m:
                rows.append([m.group(i) for i in range(1, 7)])
            elif line.startswith("Final "):
                k, v = line.split(":", 1)
                final[k.strip()] = v.strip()
        out[name] = {"params": params, "rows": rows, "final": final}
    with open(os.path.join(here, "qa_ref.json"), "w") as f:
        json.dump(out, f, indent=1)
    copies = {
        "QA/TEST_1_DHIT_DNS/DATA/vel_32x32x32.dat": "vel_32x32x32.dat",
        "QA/TEST_2_DHIT_SMAGORINSKY/DATA/vel_48x48x48.bin": "vel_48x48x48.bin",
        "QA/TEST_3_FORCING_ENERGY_CASCADE/DATA/Forced_Energy_Cascade.dat": "Forced_Energy_Cascade_K291.dat",
        "QA/TEST_9_CBC/DATA/Forced_Energy_Cascade.dat": "Forced_Energy_Cascade_K150.dat",
    }
    for src, dst in copies.items():
        d = os.path.join(here, dst)
        if not os.path.exists(d):
            shutil.copy(os.path.join(ref_root, src), d)


if __name__ == "__main__":
    main(*sys.argv[1:])
