"""The bench's driver-facing contract: one JSON line with the agreed keys
(metric, value, unit, n_gpus, steps, warmup, ms_per_step, higher_is_better,
scaling, vs_baseline, dtype, data, config with its workload, roofline and
cpu_baseline), and a clear refusal of more than one GPU."""
import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    """--gpus above 1 exits before touching a device, with a message that
    says one GPU is measured."""
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


@pytest.mark.gpu
@pytest.mark.timeout(330)
def test_json_line_contract():
    """A short run of the headline leg (the served 256^2 QU FP32 batch of 128,
    its direct and plain forms, the per-kernel timing pass and a bounded CPU
    baseline sample) prints the contract's JSON line."""
    print("[case] bench.py short run", flush=True)
    cmd = [sys.executable, str(ROOT / "bench.py"), "--steps", "2", "--warmup", "1",
           "--plain-requests", "2", "--cpu-budget", "1", "--no-serve", "--no-wf", "--no-grad",
           "--no-hmc", "--no-sim", "--no-config0", "--no-map-joint"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=ROOT,
                       env=dict(os.environ))
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith('{"metric"')]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["unit"] == "maps/s" and d["higher_is_better"] is True and d["dtype"] == "f32"
    assert d["value"] > 0 and d["ms_per_step"] > 0
    assert d["config"]["workload"] == "lenseflow_fwd_QU_256_2arcmin_n7_fp32"
    assert d["config"]["batch"] == 128
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0 < rf["frac"] <= 1 and abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] in ("port", "reference")
    assert d["rel_l2_vs_fp64_oracle"] < 1e-5
