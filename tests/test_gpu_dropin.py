"""The C++ drop-in used as the reference's main.cpp uses PathTracer: the
reference's own Cornell box scene (multi-material sphere, sphere light, floor
helpers), progressive sample batches with the finish callback, and a
Scene::CheckIntersection probe."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dropin_cornell_box_renders():
    exe = os.path.join(ROOT, "omochi-renderer3_amd", "lib", "demo_cornell")
    assert os.path.exists(exe), "build the library first"
    print("running", exe, flush=True)
    p = subprocess.run([exe, "64", "36", "3"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    summary = json.loads(p.stdout.strip().splitlines()[-1])
    assert summary["batches"] == 3 and summary["probe_hit"] == 1
    # the probe ray from (50,40,200) along -z meets the back wall plane z = 0
    assert abs(summary["probe_t"] - 200.0) < 1e-3 or summary["probe_t"] < 200.0
    assert 0.01 < summary["mean"] < 50.0
