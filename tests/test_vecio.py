"""CPU tests of the library's vector-file I/O (host code, no GPU call) and
of the command-line programs' argument handling: files written by the
library and by the compiled reference's own vector.c are byte-identical and
read back the same (doc/file_format.rst)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "yael_amd", "_lib", "bin")
f32p = ctypes.POINTER(ctypes.c_float)
i32p = ctypes.POINTER(ctypes.c_int)


@pytest.fixture(scope="module")
def lib():
    import yael_amd
    L = yael_amd.lib()
    for name in ("fvecs_fsize", "ivecs_fsize", "bvecs_fsize"):
        getattr(L, name).restype = ctypes.c_long
    L.fvec_purge_nans.restype = ctypes.c_long
    L.fvec_purge_nans.argtypes = [f32p, ctypes.c_long, ctypes.c_float]
    return L


@pytest.fixture(scope="module")
def ref():
    from oracle_bind import Reference, have_ref
    if not have_ref():
        pytest.skip("compiled reference not built")
    return Reference().lib


def _fp(a):
    return a.ctypes.data_as(f32p)


def _ip(a):
    return a.ctypes.data_as(i32p)


def test_fvecs_ivecs_bytes_match_reference(lib, ref, tmp_path):
    rng = np.random.default_rng(0)
    v = rng.standard_normal((37, 13)).astype(np.float32)
    iv = rng.integers(-1000, 1000, (29, 7)).astype(np.int32)
    for name, fn, arr, d, n, cv in (("f", "fvecs_write", v, 13, 37, _fp), ("i", "ivecs_write", iv, 7, 29, _ip)):
        a, b = tmp_path / (name + "_lib"), tmp_path / (name + "_ref")
        assert getattr(lib, fn)(str(a).encode(), d, n, cv(arr)) == n
        assert getattr(ref, fn)(str(b).encode(), d, n, cv(arr)) == n
        assert a.read_bytes() == b.read_bytes()
    # text: "%f " / "%d " per value, a newline per vector (vector.c:1279-1295,
    # 1475-1492; the reference's writers never close the file, so their
    # bytes are compared with the format itself)
    for fn, arr, d, n, cv, f in (("fvecs_write_txt", v, 13, 37, _fp, "%f "),
                                 ("ivecs_write_txt", iv, 7, 29, _ip, "%d ")):
        a = tmp_path / (fn + "_lib")
        assert getattr(lib, fn)(str(a).encode(), d, n, cv(arr)) == n
        want = "".join("".join(f % x for x in row.tolist()) + "\n" for row in arr)
        assert a.read_bytes() == want.encode()


def test_read_back_and_sizes(lib, ref, tmp_path):
    rng = np.random.default_rng(1)
    v = rng.standard_normal((50, 9)).astype(np.float32)
    p = tmp_path / "v.fvecs"
    ref.fvecs_write(str(p).encode(), 9, 50, _fp(v))
    d, n = ctypes.c_int(), ctypes.c_int()
    assert lib.fvecs_fsize(str(p).encode(), ctypes.byref(d), ctypes.byref(n)) == 50 * (9 * 4 + 4)
    assert (d.value, n.value) == (9, 50)
    out = np.empty((50, 9), np.float32)
    assert lib.fvecs_read(str(p).encode(), 9, 50, _fp(out)) == 50
    assert np.array_equal(out, v)
    # a short read: the number of vectors present
    out2 = np.empty((60, 9), np.float32)
    assert lib.fvecs_read(str(p).encode(), 9, 60, _fp(out2)) == 50
    # fvecs_new_read
    fp = ctypes.POINTER(ctypes.c_float)()
    assert lib.fvecs_new_read(str(p).encode(), ctypes.byref(d), ctypes.byref(fp)) == 50
    assert np.array_equal(np.ctypeslib.as_array(fp, shape=(50, 9)), v)
    ctypes.CDLL(None).free(fp)
    # text round trip through the reference's reader
    t = tmp_path / "v.txt"
    lib.fvecs_write_txt(str(t).encode(), 9, 50, _fp(v))
    back = np.empty((50, 9), np.float32)
    assert ref.fvecs_read_txt(str(t).encode(), 9, 50, _fp(back)) == 50
    mine = np.empty((50, 9), np.float32)
    assert lib.fvecs_read_txt(str(t).encode(), 9, 50, _fp(mine)) == 50
    assert np.array_equal(back, mine)
    # a file whose size is not a whole number of records
    bad = tmp_path / "bad.fvecs"
    bad.write_bytes(p.read_bytes()[:-3])
    assert lib.fvecs_fsize(str(bad).encode(), ctypes.byref(d), ctypes.byref(n)) == -1


def test_bvecs_as_floats(lib, tmp_path):
    rng = np.random.default_rng(2)
    b = rng.integers(0, 256, (20, 16)).astype(np.uint8)
    p = tmp_path / "b.bvecs"
    with open(p, "wb") as f:
        for row in b:
            f.write(np.int32(16).tobytes() + row.tobytes())
    d, n = ctypes.c_int(), ctypes.c_int()
    assert lib.bvecs_fsize(str(p).encode(), ctypes.byref(d), ctypes.byref(n)) > 0
    assert (d.value, n.value) == (16, 20)
    out = np.empty((20, 16), np.float32)
    assert lib.b2fvecs_read(str(p).encode(), 16, 20, _fp(out)) == 20
    assert np.array_equal(out, b.astype(np.float32))


def test_purge_nans(lib):
    v = np.array([1, np.nan, 3, np.nan], np.float32)
    assert lib.fvec_purge_nans(_fp(v), 4, 2.0) == 2
    assert np.array_equal(v, [1, 2, 3, 2])


def test_cli_usage_and_missing_arguments():
    for prog in ("kmeans", "knn"):
        r = subprocess.run([os.path.join(BIN, prog), "-h"], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0 and r.stdout.startswith("Usage:")
        r = subprocess.run([os.path.join(BIN, prog), "-k", "3"], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode != 0 and "required" in r.stderr


def test_malformed_files_same_results_as_reference(lib, ref, tmp_path):
    """A payload cut short, a trailing fragment shorter than a length word,
    a record of another length and an empty file: sizes, bounded reads and
    allocating reads return what the reference's own readers return."""
    for L in (lib, ref):
        L.fvecs_fsize.restype = ctypes.c_long
    v = np.arange(40, dtype=np.float32).reshape(8, 5)
    good = b"".join(np.int32(5).tobytes() + r.tobytes() for r in v)
    cases = {
        "good": good,
        "cut_payload": good[:-3],
        "cut_len": good + b"\x05\x00",
        "baddim": good[:72] + np.int32(4).tobytes() + good[76:],
        "empty": b"",
    }
    for name, data in cases.items():
        p = tmp_path / (name + ".fvecs")
        p.write_bytes(data)
        got = []
        for L in (lib, ref):
            d, n = ctypes.c_int(), ctypes.c_int()
            size = L.fvecs_fsize(str(p).encode(), ctypes.byref(d), ctypes.byref(n))
            out = np.zeros((10, 5), np.float32)
            r = L.fvecs_read(str(p).encode(), 5, 10, _fp(out))
            res = [size, d.value, n.value, r, out[:max(r, 0)].tobytes()]
            if name != "empty":  # the reference's allocating reader is not defined on it
                fp, d2 = f32p(), ctypes.c_int()
                r2 = L.fvecs_new_read(str(p).encode(), ctypes.byref(d2), ctypes.byref(fp))
                res += [r2, d2.value]
                if r2 > 0:
                    res.append(np.ctypeslib.as_array(fp, shape=(r2, d2.value)).tobytes())
                    ctypes.CDLL(None).free(fp)
            got.append(res)
        assert got[0] == got[1], name
