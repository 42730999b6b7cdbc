"""bench.py on the GPU, at a small size: the JSON line keeps its contract
(the fields, the roofline and cpu_baseline objects, the placement trials)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_json_line(need_gpu):
    args = ["--steps", "3", "--warmup", "1", "--size", "128", "--variants", "bf16",
            "--accuracy-steps", "0", "--adjoint-steps", "0", "--cpu-seconds", "1",
            "--workloads", "", "--placement-trials", "2"]
    print("bench.py", " ".join(args), flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["unit"] == "MLUPS" and d["value"] > 0 and d["higher_is_better"] is True
    assert d["dtype"] == "f32" and d["config"]["workload"] == "zsc-128^3"
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0 < rf["frac"] < 1 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] in ("port", "reference")
    assert len(d["placement_ms"]) == 2 and min(d["placement_ms"]) > 0 and d["placement_s"] > 0
    lo, hi = rf["frac_placement_range"]
    assert lo <= rf["frac_placement_median"] <= hi
    # copy ceilings on the timed lattice pair, and the step's share of them
    for k in ("linear", "march"):
        assert rf[f"copy_{k}_GBs"] > 0
        assert abs(rf[f"frac_of_copy_{k}"] - rf["achieved"] / rf[f"copy_{k}_GBs"]) < 1e-3
    assert cb["cpu_model"] and cb["fp32"]["value"] > 0
    v = d["variants"]["bf16"]
    assert v["mlups"] > 0 and 0 < v["roofline"]["frac"] < 1
    assert "fp16" in v["storage"] and "bf16" in v["storage"]
    assert v["roofline"]["copy_march_GBs"] > 0
