"""CPU model of the discovery protocol and the barriers (oracle/oracle.c,
orc_dp_model): groups are threads, at most `resident` run at once, and a
group only gives up its slot when it finishes (occupancy-bound execution).
This is where deadlock is forced on purpose: never on the GPU."""
import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("lock", [0, 1, 2])
@pytest.mark.parametrize("bar", [0, 1, 2])
def test_discovery_finds_only_coresident_groups(lock, bar):
    r = oracle.dp_model(groups=48, resident=12, lock=lock, bar=bar, barriers=12, watchdog_ms=3000)
    assert not r["deadlock"]
    P = r["P"]
    assert 1 <= P <= 12
    ids = r["ids"][r["ids"] >= 0]
    assert sorted(ids.tolist()) == list(range(P))
    assert r["violations"] == 0


def test_all_resident_all_participate_lockfree():
    r = oracle.dp_model(groups=16, resident=16, lock=0, bar=1, barriers=30, watchdog_ms=3000)
    assert not r["deadlock"] and r["P"] == 16 and r["violations"] == 0


@pytest.mark.parametrize("bar", [0, 1, 2])
def test_skip_mode_beyond_residency_deadlocks(bar):
    # Non-portable barrier (skip mode): more launched than resident groups.
    r = oracle.dp_model(groups=24, resident=8, lock=0, bar=bar, skip=1, barriers=3, watchdog_ms=400)
    assert r["deadlock"]


def test_skip_mode_within_residency_completes():
    r = oracle.dp_model(groups=8, resident=8, lock=0, bar=2, skip=1, barriers=20, watchdog_ms=3000)
    assert not r["deadlock"] and r["P"] == 8 and r["violations"] == 0


# ---- the model under ThreadSanitizer ---------------------------------------
# The model carries the device's memory orders (relaxed polls, acq_rel
# read-modify-writes, release publication) and a plain payload written
# before each barrier and read after it.  A ThreadSanitizer build must find
# no race with those orders, and must find one when the barrier's arrival
# and release are planted relaxed (the mistake a wrong fence would be).
import os  # noqa: E402
import subprocess  # noqa: E402
from pathlib import Path  # noqa: E402

TSAN = Path(oracle.__file__).resolve().parent / "_tsan" / "dp_model_tsan"


@pytest.fixture(scope="module")
def tsan_model():
    subprocess.run(["make", "-s", "-C", str(TSAN.parents[1]), "tsan"], check=True)
    return TSAN


def run_tsan(exe, groups, resident, lock, bar, skip, barriers, weak):
    env = {**os.environ, "TSAN_OPTIONS": "halt_on_error=0 report_signal_unsafe=0"}
    p = subprocess.run([str(exe), *map(str, (groups, resident, lock, bar, skip, barriers, weak))],
                       capture_output=True, text=True, timeout=240, env=env)
    return p.stdout + p.stderr


@pytest.mark.parametrize("lock", [0, 1, 2])
@pytest.mark.parametrize("bar", [0, 1, 2])
def test_model_is_race_free_under_tsan(tsan_model, lock, bar):
    out = run_tsan(tsan_model, 20, 8, lock, bar, 0, 25, 0)
    assert "ThreadSanitizer" not in out, out[-3000:]
    assert "violations=0 deadlock=0" in out


@pytest.mark.parametrize("bar", [0, 1, 2])
def test_tsan_catches_a_relaxed_barrier(tsan_model, bar):
    out = run_tsan(tsan_model, 16, 16, 0, bar, 0, 25, 1)
    assert "WARNING: ThreadSanitizer: data race" in out
