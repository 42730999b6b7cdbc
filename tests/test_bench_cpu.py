"""bench.py contract pieces that need no GPU."""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in out["error"] and out["n_gpus"] == 2


def test_traffic_summary_is_read_and_at_least_algorithmic():
    import bench
    t = bench.traffic_from_profile((512, 512, 512))
    cells = 512 ** 3
    assert set(t) == {"f32", "bf16", "fp16", "f32d"}
    assert t["f32"] >= 208 * cells and t["bf16"] >= 108 * cells and t["fp16"] >= 108 * cells
    assert t["f32d"] >= 208 * cells
    assert bench.traffic_from_profile((7, 7, 7)) == {}


def test_z_slab_is_one_share_of_the_box():
    from lbm_bench_gpu_amd import configs
    full = configs.workload(3)
    seen = 0
    for r in range(8):
        w = configs.z_slab(3, 8, r)
        assert w.dims == (1024, 1024, 128)
        assert w.params().values == full.params().values
        b = w.bubbles
        assert (b[:, 2] - b[:, 3] - 4.0 >= 1).all() and (b[:, 2] + b[:, 3] + 4.0 <= 128).all()
        seen += len(b)
    assert 0 < seen <= len(full.bubbles)
    with pytest.raises(ValueError):
        configs.z_slab(3, 7)
