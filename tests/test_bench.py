"""bench.py contract pieces that run without a GPU: more than one GPU is
refused with a message, and the CPU baseline reports the fields the bench
line carries."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_baseline_fields():
    sys.path.insert(0, ROOT)
    import bench
    from spectrum_amd import capi

    x = capi.siggen(capi.GEN_MULTI_TONE, 0x5EED0002, 3, 4096)
    out = bench.cpu_baseline(x, 4096, 64, target_s=5.0)
    assert set(out) >= {"value", "unit", "cores", "kind", "sample"}
    assert out["unit"] == "channels/s" and out["cores"] == 1 and out["kind"] in ("reference", "port")
    assert np.isfinite(out["value"]) and out["value"] > 0
