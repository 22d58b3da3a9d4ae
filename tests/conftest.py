import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

# the compiled reference links OpenBLAS: keep it single-threaded inside the
# OpenMP workers, as yael's README recommends for its BLAS
os.environ["OPENBLAS_NUM_THREADS"] = "1"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on cuda:0)")
