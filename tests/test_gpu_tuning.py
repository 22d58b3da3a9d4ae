"""GPU test of the launch-plan autotuner: the FP32 and precision-form
searches time their candidate plans on the first call of a shape class and
record the fastest in the tuning cache; every plan gives the same bits, so
a process that tunes and one that runs the default plans (YAEL_MI355X_TUNE=0)
return identical results."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import hashlib, sys
sys.path.insert(0, %(root)r)
import numpy as np
import yael_amd
from yael_amd import device as dev
from yael_amd.workloads import mixture
yael_amd.init()
b = mixture(5, 200000, 128, 512, 20, centres_seed=0)
q = mixture(6, 3000, 128, 512, 20, centres_seed=0)
B, Q = dev.DeviceSet(b), dev.DeviceSet(q)
h = hashlib.sha256()
for k in (1, 10, 100):
    for rep in range(2):  # first call tunes (when enabled), second uses the plan
        i, d = dev.knn_fp32(Q, B, k)
        h.update(i.tobytes()); h.update(d.tobytes())
        i, d, _ = dev.knn_bf16(Q, B, k, max(k, 8) if k <= 16 else 2 * k)
        h.update(i.tobytes()); h.update(d.tobytes())
print("DIGEST", h.hexdigest(), flush=True)
'''


def run(env):
    r = subprocess.run([sys.executable, "-c", SCRIPT % {"root": ROOT}], env=env,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-3000:]
    line = [x for x in r.stdout.splitlines() if x.startswith("DIGEST")]
    assert line, r.stdout[-2000:]
    return line[0].split()[1]


def test_tuned_plans_give_the_default_plans_bits(tmp_path):
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, YAEL_MI355X_TUNE_CACHE=str(cache))
    print("tuning run", flush=True)
    tuned = run(env)
    print("default-plan run", flush=True)
    plain = run(dict(env, YAEL_MI355X_TUNE="0"))
    assert tuned == plain
    lines = cache.read_text().split("\n")
    print("\n".join(lines), flush=True)
    # shape classes with more than one distinct candidate plan were tuned
    # (the others run their only plan) and recorded with a plan index
    # (plan indices: 3 work-item targets x 4 sketch-stride scales in both
    # forms)
    rec = [x.split() for x in lines if x.strip()]
    nplans = {"f32": 12, "bf16": 12}
    assert all(len(r) == 2 and 0 <= int(r[1]) < nplans.get(r[0].split(":")[0], 12) for r in rec)
    assert sum(r[0].startswith("f32:") for r in rec) >= 2
    assert sum(r[0].startswith("bf16:") for r in rec) >= 1
