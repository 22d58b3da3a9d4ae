"""GPU tests of the command-line programs (yael's progs/kmeans.c and
progs/knn.c over the library): on integer-valued data their output files
hold the oracle's results bit for bit."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "yael_amd", "_lib", "bin")


@pytest.fixture(scope="module")
def oracle():
    from oracle_bind import Oracle
    o = Oracle()
    o.set_blas(0, 0)
    return o


def write_vecs(path, a, kind):
    a = np.ascontiguousarray(a)
    with open(path, "wb") as f:
        for row in a:
            f.write(np.int32(a.shape[1]).tobytes() + row.astype(kind).tobytes())


def read_vecs(path, kind):
    raw = np.fromfile(path, dtype=np.uint8)
    d = int(raw[:4].view(np.int32)[0])
    rec = raw.reshape(-1, 4 + 4 * d)
    assert np.all(rec[:, :4].copy().view(np.int32) == d)
    return rec[:, 4:].copy().view(kind)


def run(args):
    print("run", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def mixture(rng, n, d, ncomp=20):
    cen = rng.uniform(0, 255, (ncomp, d))
    return np.clip(np.rint(cen[rng.integers(0, ncomp, n)] + 20 * rng.standard_normal((n, d))),
                   0, 255).astype(np.float32)


def test_kmeans_cli(oracle, tmp_path):
    rng = np.random.default_rng(5)
    v = mixture(rng, 3000, 16)
    src = tmp_path / "v.fvecs"
    write_vecs(src, v, np.float32)
    out, outt = tmp_path / "c.fvecs", tmp_path / "c.txt"
    run([os.path.join(BIN, "kmeans"), "-i", str(src), "-k", "20", "-niter", "1", "-seed", "7",
         "-o", str(out)])
    run([os.path.join(BIN, "kmeans"), "-i", str(src), "-k", "20", "-niter", "1", "-seed", "7",
         "-otxt", str(outt)])
    _, cent, *_ = oracle.kmeans(v, 20, 1, flags=1, seed=7)
    got = read_vecs(out, np.float32)
    assert np.array_equal(got.view(np.int32), cent.view(np.int32))
    np.testing.assert_allclose(np.loadtxt(outt), cent, atol=5e-7 * 255)


def test_knn_cli(oracle, tmp_path):
    rng = np.random.default_rng(6)
    b = mixture(rng, 5000, 24)
    q = mixture(rng, 200, 24)
    fb, fq, fqb = tmp_path / "b.fvecs", tmp_path / "q.fvecs", tmp_path / "q.bvecs"
    write_vecs(fb, b, np.float32)
    write_vecs(fq, q, np.float32)
    with open(fqb, "wb") as f:
        for row in q.astype(np.uint8):
            f.write(np.int32(24).tobytes() + row.tobytes())
    nn, dis = tmp_path / "nn.ivecs", tmp_path / "dis.fvecs"
    run([os.path.join(BIN, "knn"), "-b", str(fb), "-q", str(fq), "-k", "7", "-onn", str(nn),
         "-odis", str(dis)])
    a, d = oracle.knn_full(2, b, q, 7)
    a, d = oracle.knn_reorder_shortlist(b, q, a, d)
    assert np.array_equal(read_vecs(nn, np.int32), a)
    assert np.array_equal(read_vecs(dis, np.float32).view(np.int32), d.view(np.int32))
    # bvecs queries, text outputs
    nnt, dist = tmp_path / "nn.txt", tmp_path / "dis.txt"
    run([os.path.join(BIN, "knn"), "-b", str(fb), "-qb", str(fqb), "-k", "7", "-onnt", str(nnt),
         "-odist", str(dist), "-silence"])
    assert np.array_equal(np.loadtxt(nnt, dtype=np.int64), a)
    np.testing.assert_allclose(np.loadtxt(dist), d, rtol=1e-6)
