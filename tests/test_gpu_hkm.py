"""GPU tests of the hierarchical k-means (hkm.c): every level one batched
k-means on the GPU, node seeds from lrand48() in node order, points of a
node in (label, index) order; the descent kernel for hkm_quantize."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


@pytest.fixture(scope="module")
def hk():
    import yael_amd
    from yael_amd import hkm
    yael_amd.init()
    return hkm


def srand48(seed):
    import ctypes
    ctypes.CDLL(None).srand48(ctypes.c_long(seed))


def test_learn_one_step_per_node_bitexact(hk):
    """One k-means step per node on integer data: exact arithmetic, so the
    tree, the labels and the file are the reference's bits."""
    g = load("hkm_step")
    srand48(int(g["srand48"]))
    print("hkm_learn n=6000 d=16 nlevel=3 bf=4", flush=True)
    t, labels = hk.hkm_learn(g["v"], 3, 4, niter=1)
    for l in range(3):
        assert np.array_equal(t.centroids(l).view(np.int32), g["c%d" % l].view(np.int32)), l
    assert np.array_equal(labels, g["labels"])


def test_quantize_integer_tree_bitexact(hk, tmp_path):
    """Descent through a tree of integer centroids (exact distances; ties to
    the lowest child)."""
    g = load("hkm_quant")
    bf, d = int(g["bf"]), g["c0"].shape[1]
    path = tmp_path / "t.hkm"
    with open(path, "wb") as f:
        f.write(np.array([3, bf, d], np.int32).tobytes())
        for l in range(3):
            c = np.ascontiguousarray(g["c%d" % l], np.float32)
            f.write(np.array([c.size], np.int32).tobytes() + c.tobytes())
    t = hk.hkm_read(path)
    print("hkm_quantize 2000 points", flush=True)
    assert np.array_equal(t.quantize(g["q"]), g["qidx"])


def test_quantize_learned_tree(hk):
    """The learned tree's centroids are means (not integers): the descent
    agrees with the reference up to FP32 near-ties."""
    g = load("hkm_step")
    srand48(int(g["srand48"]))
    t, _ = hk.hkm_learn(g["v"], 3, 4, niter=1)
    idx = t.quantize(g["q"])
    assert (idx == g["qidx"]).mean() >= 0.99


def test_learn_float_data(hk):
    """Float data, six iterations per node: the same seeds; trajectories may
    part only at near-ties."""
    g = load("hkm_float")
    srand48(int(g["srand48"]))
    t, labels = hk.hkm_learn(g["v"], 2, 5, niter=6)
    assert (labels == g["labels"]).mean() > 0.95
    np.testing.assert_allclose(t.centroids(0), g["c0"], rtol=1e-3, atol=1e-3)


def test_learn_write_read_roundtrip(hk, tmp_path):
    g = load("hkm_step")
    srand48(1)
    t, labels = hk.hkm_learn(g["v"], 2, 6, niter=4)
    assert labels.min() >= 0 and labels.max() < 36
    p = tmp_path / "rt.hkm"
    t.write(p)
    r = hk.hkm_read(p)
    for l in range(2):
        assert np.array_equal(r.centroids(l), t.centroids(l))
    assert np.array_equal(r.quantize(g["q"]), t.quantize(g["q"]))


def test_learn_rejects_small_node(hk):
    """A node with fewer points than bf: an error (the reference asserts)."""
    from yael_amd import YaelGPUError
    v = np.repeat(np.arange(3, dtype=np.float32)[:, None], 4, axis=1)
    v = np.concatenate([v, v + 100])
    with pytest.raises(YaelGPUError):
        hk.hkm_learn(v, 3, 4, niter=2)


def _tree_file(path, levels, bf):
    """A tree in the reference's file format (hkm.c:181-199): nlevel, bf, d,
    then per level the element count and the centroids."""
    d = levels[0].shape[1]
    with open(path, "wb") as f:
        f.write(np.array([len(levels), bf, d], np.int32).tobytes())
        for c in levels:
            c = np.ascontiguousarray(c, np.float32)
            f.write(np.array([c.size], np.int32).tobytes() + c.tobytes())


@pytest.mark.parametrize("nlevel,bf,d,n", [(3, 10, 40, 50000), (2, 32, 128, 200000),
                                           (2, 45, 20, 30011)])
def test_quantize_levels_batched_vs_oracle(hk, tmp_path, nlevel, bf, d, n):
    """The level-at-a-time descent (points ordered by node, each node one
    instance of the batched FP32 assign) on integer-valued trees and points
    (every distance exact): the oracle's leaves bit for bit, with branching
    factors on and off the 32-row block size."""
    from oracle_bind import Oracle
    rng = np.random.default_rng(nlevel * 1000 + bf)
    levels = [np.rint(rng.uniform(0, 60, (bf ** (l + 1), d))).astype(np.float32)
              for l in range(nlevel)]
    v = np.rint(rng.uniform(0, 60, (n, d))).astype(np.float32)
    path = tmp_path / "int.hkm"
    _tree_file(path, levels, bf)
    t = hk.hkm_read(path)
    print("hkm_quantize nlevel=%d bf=%d d=%d n=%d" % (nlevel, bf, d, n), flush=True)
    got = t.quantize(v)
    want = Oracle().hkm_quantize(levels, bf, v)
    assert np.array_equal(got, want)
    assert got.min() >= 0 and got.max() < bf ** nlevel
