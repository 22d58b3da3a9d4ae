"""CPU checks of the C-ABI library: it loads, exports every function that
include/yael_mi355x.h and yael_amd/csrc/yk.h declare, and fails loudly
(an error code and message, never a CPU fallback) where no GPU is visible.
No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "yael_mi355x.h"),
           os.path.join(ROOT, "yael_amd", "csrc", "yk.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
        src = re.sub(r"#[^\n]*", " ", src)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", src):
            name = m.group(1)
            if name not in ("if", "sizeof", "defined"):
                names.add(name)
    return sorted(names)


def test_library_built_and_loads():
    import yael_amd
    assert os.path.exists(yael_amd.LIB_PATH), "run __graft_entry__.build()"
    yael_amd.lib()


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    import yael_amd
    assert hasattr(yael_amd.lib(), name), name


def test_reference_symbols_are_c_abi():
    """The reference's entry points are plain C symbols (no C++ mangling)."""
    import yael_amd
    out = subprocess.run(["nm", "-D", "--defined-only", yael_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in ("kmeans", "knn_full", "knn_full_thread", "compute_cross_distances", "nn", "knn",
              "knn_reorder_shortlist"):
        assert s in syms, s


def _has_gpu():
    return os.path.exists("/dev/kfd") and subprocess.run(
        ["bash", "-c", "ls /dev/dri/renderD* >/dev/null 2>&1"]).returncode == 0


def test_no_gpu_fails_loudly():
    """Without a GPU the library reports an error; it never computes on the CPU."""
    if _has_gpu():
        pytest.skip("a GPU is visible")
    code = ("import numpy as np, yael_amd\n"
            "try:\n"
            "    yael_amd.knn(np.zeros((2,4),np.float32), np.zeros((3,4),np.float32), 1)\n"
            "except yael_amd.YaelGPUError as e:\n"
            "    print('RAISED', e)\n")
    r = subprocess.run(["python3", "-c", code], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    assert "RAISED" in r.stdout, r.stdout + r.stderr


def test_oracle_not_linked_by_product():
    import yael_amd
    out = subprocess.run(["ldd", yael_amd.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "yael_ref" not in out
    strings = open(yael_amd.LIB_PATH, "rb").read()
    assert b"liboracle" not in strings and b"oy_knn_full" not in strings
