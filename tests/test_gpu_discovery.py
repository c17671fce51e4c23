"""Discovery protocol and barriers on the GPU (through the C ABI).

Invariants (the reference holds no golden vectors for this part): discovered
ids form 0..P-1 with 1 <= P <= launched and P <= the residency bound; skip
mode reports every launched group; after each barrier k the per-round
arrival count equals P*k and plain stores made before the barrier are
visible after it; the context resets itself between launches."""
import pytest

import gpu_discovery_barrier_amd as gdb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


def test_device_query():
    txt = gdb.device_query()
    print(txt)
    assert "gfx950" in txt


@pytest.mark.parametrize("wgs", [64, 256, 1024])
def test_occupancy_bound_positive(wgs):
    bound = gdb.max_resident_groups(wgs, 0)
    print("wgs", wgs, "resident bound", bound)
    assert bound >= 256


@pytest.mark.parametrize("lock", [gdb.PROTOCOL_LOCKFREE, gdb.PROTOCOL_TICKET, gdb.PROTOCOL_SPIN])
@pytest.mark.parametrize("wgs", [64, 256])
def test_discovery_within_bound(ctx, lock, wgs):
    ctx.configure(lock, gdb.BARRIER_COUNTER)
    bound = gdb.max_resident_groups(wgs, 0)
    for wgc in (1, 7, min(bound, 512), bound):
        p = ctx.occupancy_test(wgc, wgs, 0, prot=True)
        print("lock", lock, "wgs", wgs, "launched", wgc, "P", p)
        assert 1 <= p <= wgc
    ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)


def test_discovery_over_launch_is_safe(ctx):
    # Launch 4x the residency bound: only co-resident groups may participate.
    wgs = 256
    bound = gdb.max_resident_groups(wgs, 0)
    p = ctx.occupancy_test(4 * bound, wgs, 0, prot=True)
    print("over-launch", 4 * bound, "P", p, "bound", bound)
    assert 1 <= p <= bound


def test_skip_mode_within_bound(ctx):
    wgs = 256
    bound = gdb.max_resident_groups(wgs, 0)
    p = ctx.occupancy_test(bound, wgs, 0, prot=False)
    assert p == bound


def test_skip_mode_beyond_bound_is_refused(ctx):
    wgs = 256
    bound = gdb.max_resident_groups(wgs, 0)
    with pytest.raises(gdb.GdbError) as e:
        ctx.occupancy_test(bound + 1, wgs, 0, prot=False)
    assert e.value.code == gdb.GDB_E_UNSAFE


@pytest.mark.parametrize("bar", [gdb.BARRIER_XF, gdb.BARRIER_COUNTER, gdb.BARRIER_TREE, gdb.BARRIER_AUTO])
@pytest.mark.parametrize("wgs,groups", [(64, 3), (256, 64), (256, 1024), (1024, 256), (64, 4096)])
def test_barrier_invariants(ctx, bar, wgs, groups):
    bound = gdb.max_resident_groups(wgs, 0)
    groups = min(groups, bound)
    ctx.configure(gdb.PROTOCOL_LOCKFREE, bar)
    bad, p = ctx.barrier_check(wgs, groups, 200)
    print("bar", bar, "wgs", wgs, "groups", groups, "P", p, "violations", bad)
    assert bad == 0
    assert 1 <= p <= groups
    ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)


@pytest.mark.parametrize("xcd", [0, 3])
@pytest.mark.parametrize("wgs,groups", [(256, 64), (1024, 64), (256, 2048)])
def test_single_xcd_barrier_invariants(ctx, xcd, wgs, groups):
    groups = min(groups, gdb.max_resident_groups(wgs, 0))
    ctx.set_xcd_only(xcd)
    try:
        bad, p = ctx.barrier_check(wgs, groups, 300)
        print("xcd", xcd, "wgs", wgs, "groups", groups, "P", p, "violations", bad)
        assert bad == 0
        assert 1 <= p <= (groups + 7) // 8 + 1
    finally:
        ctx.set_xcd_only(-1)


def test_tree_barrier_repeated_launches(ctx):
    # Membership counts live in a buffer chosen by launch parity; several
    # launches in a row must each see exact counts.
    ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_TREE)
    try:
        for groups in (512, 128, 2048, 77):
            bad, p = ctx.barrier_check(256, groups, 50)
            assert bad == 0 and p == groups
    finally:
        ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)


def test_context_self_resets_between_launches(ctx):
    ctx.init()
    ps = [ctx.barrier_bench(256, 64, 10)[1] for _ in range(3)]
    assert all(p == ps[0] for p in ps)


def test_timeout_ends_wait_instead_of_hanging(ctx):
    # A wait bound far below one barrier's cost forces the timeout path from
    # the host; every kernel still exits normally.
    ctx.set_timeouts(0.0, 20.0)
    ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)
    try:
        with pytest.raises(gdb.GdbError) as e:
            ctx.barrier_bench(256, 256, 2000)
        assert e.value.code == gdb.GDB_E_TIMEOUT
    finally:
        ctx.set_timeouts(10e6, 20.0)
        ctx.init()
    ms, p = ctx.barrier_bench(256, 64, 10)
    assert p >= 1


@pytest.mark.parametrize("bar", [gdb.BARRIER_XF, gdb.BARRIER_COUNTER, gdb.BARRIER_TREE])
def test_solve_after_forced_timeout_matches_oracle(bar):
    # A zero wait bound (set from the host) makes the first barrier give up
    # part-way, leaving arrival counts / flags behind.  With the bound
    # restored and no explicit init, the next solve on the same context must
    # still be exact: the library resets the context when it reports the
    # timeout.
    import numpy as np
    import oracle

    g = gdb.gen_grid(96, 96, 0.05, seed=2, wlo=1, whi=255)
    dg = gdb.DeviceCSR(g)
    want = oracle.dijkstra_tb(g.row, g.col, g.w, 0)
    with gdb.Context(bar_kind=bar) as c:
        o = gdb.Opts.make(variant=gdb.VARIANT_TUNED, wgs=256, groups=96)
        c.set_timeouts(0.0, 20.0)
        with pytest.raises(gdb.GdbError) as e:
            gdb.sssp_worklist(c, dg, 0, o)
        assert e.value.code == gdb.GDB_E_TIMEOUT
        c.set_timeouts(10e6, 20.0)
        for _ in range(2):
            got, st = gdb.sssp_worklist(c, dg, 0, o)
            assert (got.astype(np.int64) == want).all()
            assert st.participating > 1


def test_time_protocol(ctx):
    ns = ctx.time_protocol(1000, 256)
    print("protocol ns", ns, "P", ctx.participating())
    assert ns > 0


def test_occupancy_suite():
    # The counterpart of run_occupancy_tests.py / run_occupancy_timing.py,
    # two runs per configuration and lock kind.
    from gpu_discovery_barrier_amd import occupancy_suite

    res = occupancy_suite.run(iterations=2, groups=1000)
    text = occupancy_suite.report(res)
    print(text)
    assert len(res["configs"]) == 4
    for d in res["configs"]:
        assert d["occupancy"] > 0 and d["skip_mode_at_bound_completes"] and d["one_more_refused"]
        for name, _ in occupancy_suite.LOCKS:
            r = d[name]
            assert len(r["occupancy_list"]) == 2 and r["std_dev"] >= 0
            assert all(1 <= p <= min(1000, d["occupancy"]) for p in r["occupancy_list"])
            assert r["protocol_ms_average"] > 0
        # The lock-free protocol admits every co-resident group it can see.
        assert d["lock-free"]["average"] >= min(1000, d["occupancy"]) * 0.5
    assert "true occupancy:" in text and "ticket occupancy list:" in text
