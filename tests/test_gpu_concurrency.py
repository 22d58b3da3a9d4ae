"""Concurrent calls from several host threads (yael's functions are
re-entrant and its SWIG layer releases the GIL on every call, so callers may
run them in parallel; the library serialises them on one mutex).  Every
concurrent result must equal the same call made alone, bit for bit."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ya():
    import yael_amd
    yael_amd.init()
    return yael_amd


def integer_set(rng, n, d):
    return rng.integers(0, 256, (n, d)).astype(np.float32)


def test_concurrent_knn_and_kmeans_match_serial(ya):
    rng = np.random.default_rng(17)
    jobs = []
    for j in range(4):
        b = integer_set(rng, 20000 + 997 * j, 64 + 16 * j)
        q = integer_set(rng, 300 + 50 * j, b.shape[1])
        jobs.append(("knn", (q, b, 10 + j)))
    for j in range(3):
        v = integer_set(rng, 6000 + 500 * j, 32)
        jobs.append(("kmeans", (v, 16 + 8 * j, 5, 100 + j)))

    def run(job):
        kind, args = job
        if kind == "knn":
            q, b, k = args
            return ya.knn(q, b, k)
        v, k, niter, seed = args
        # a nonzero seed: seed 0 draws lrand48(), whose order across
        # threads is not defined
        c, qerr, dis, assign, nassign = ya.kmeans(v, k, niter=niter, seed=seed, verbose=False,
                                                  output="all")
        return c, np.float32(qerr), dis, assign, nassign

    serial = [run(job) for job in jobs]
    print("serial results done", flush=True)

    results = {}
    errors = []

    def worker(t):
        try:
            for rep in range(2):
                for i in range(len(jobs)):
                    j = (i + t) % len(jobs)  # threads start on different jobs
                    results[(t, rep, j)] = run(jobs[j])
        except Exception as e:  # reported below, not swallowed
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a worker thread did not finish"
    assert not errors, errors
    assert len(results) == 6 * 2 * len(jobs)
    for (t, rep, j), got in results.items():
        for a, b in zip(got, serial[j]):
            a = np.atleast_1d(np.asarray(a))
            b = np.atleast_1d(np.asarray(b))
            assert a.dtype == b.dtype and a.shape == b.shape
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), \
                "thread %d rep %d job %d differs from the serial call" % (t, rep, j)
