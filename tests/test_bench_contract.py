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


def test_headline_engine_follows_the_tuning_table(tmp_path, monkeypatch):
    # A batch_owned row (written by the autotuner only when the owned engine
    # won) makes the owned engine the headline; without it the shared row is
    # used; --engine forces either.
    from gpu_discovery_barrier_amd import autotune

    n = bench.GRID_SIDE * bench.GRID_SIDE
    rows = [autotune.row("batch", 1 << 30, 128, 256, 128, 4096, 256, 2048),
            autotune.row("batch_owned", 1 << 30, 1024, 256, 128, 8192, 0, 2048)]
    table = tmp_path / "tune_table.txt"
    autotune.write_table(rows, table)
    assert autotune.lookup("batch_owned", n, table)["delta"] == 8192
    monkeypatch.setattr(autotune, "TABLE", table)
    monkeypatch.setattr(autotune, "lookup", lambda s, nn, path=table, _f=autotune.lookup: _f(s, nn, path))
    shape, team, qcap, owned = bench.tuned_c2()
    assert owned and shape == (1024, 256, 8192, 0, 0) and team == 128 and qcap == 2048
    shape, team, qcap, owned = bench.tuned_c2("shared")
    assert not owned and shape == (128, 256, 4096, 256, 0)
    autotune.write_table(rows[:1], table)
    assert bench.tuned_c2()[3] is False
    assert bench.tuned_c2("owned")[3] is True
