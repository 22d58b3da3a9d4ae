"""Host-to-device uploads: large contiguous sets go through the staged path
(host threads copying slices through pinned buffers on their own streams),
smaller or padded sets through one copy; every byte must arrive."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ya():
    import yael_amd
    yael_amd.init()
    return yael_amd


@pytest.mark.parametrize("n,d", [(1000, 128), (131072 + 7, 128), (700001, 128), (90001, 200)])
def test_upload_roundtrip(ya, n, d):
    """(131072 + 7) x 128 and 700001 x 128 floats are 64 MB and 358 MB:
    staged, with slices and chunks that do not divide evenly; 90001 x 200
    is padded (row stride 208)."""
    from yael_amd import device
    rng = np.random.default_rng(n + d)
    x = rng.standard_normal((n, d)).astype(np.float32)
    x[-1] = np.arange(d, dtype=np.float32)  # the last row marks the tail
    print("upload n=%d d=%d (%.0f MB)" % (n, d, x.nbytes / 2**20), flush=True)
    V = device.DeviceSet(x)
    back = np.empty_like(x)
    ya.check(ya.lib().yk_download_vectors(back.ctypes.data, V.vec.ptr, n, d), "download")
    V.free()
    assert np.array_equal(back.view(np.uint32), x.view(np.uint32))
