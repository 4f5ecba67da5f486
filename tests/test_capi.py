"""The C-ABI library loads and exports every entry point include/spns.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spns.h")
LIB = os.path.join(ROOT, "spns_amd", "libspns.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spns_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ["spns_create", "spns_step", "spns_get_state", "spns_fft_c2r", "spns_fft_r2c"]:
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libspns.so not built")
def test_library_exports_all_declared_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(LIB)
    for n in declared():
        assert hasattr(lib, n)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libspns.so not built")
def test_python_binding_covers_header():
    import spns_amd
    assert set(declared()) == set(spns_amd.EXPORTS)


def test_header_compiles_as_c():
    r = subprocess.run(["gcc", "-fsyntax-only", "-x", "c", "-std=c99", HEADER], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
