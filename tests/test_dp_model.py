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
