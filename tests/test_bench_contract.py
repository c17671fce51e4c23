"""bench.py's output contract, on the CPU: the one-line JSON carries every
field the measurement consumers read, and a multi-GPU request is refused
before any GPU work."""
import json
import subprocess
import sys
import types
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

REQUIRED = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
            "scaling", "vs_baseline", "dtype", "data", "config")


def test_line_has_every_contract_field():
    args = types.SimpleNamespace(steps=5, warmup=2)
    extra = {"roofline": {"bound": "hbm", "achieved": 1.0, "peak": 8000.0, "unit": "GB/s", "frac": 1.25e-4,
                          "traffic": None},
             "cpu_baseline": {"value": 0.005, "unit": "GTEPS", "cores": 1, "kind": "port", "sample": "s"}}
    out = bench.line("GTEPS", 1.23456, "GTEPS", args, 22.4567, {"workload": "grid"}, "u32", extra)
    for k in REQUIRED:
        assert k in out, k
    assert out["n_gpus"] == 1 and out["steps"] == 5 and out["warmup"] == 2
    assert out["value"] == 1.2346 and out["ms_per_step"] == 22.457
    assert out["config"]["workload"] == "grid"
    assert set(out["roofline"]) >= {"bound", "achieved", "peak", "unit", "frac", "traffic"}
    assert set(out["cpu_baseline"]) >= {"value", "unit", "cores", "kind", "sample"}
    json.dumps(out)  # one JSON line


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""
