"""Colouring parity: every schedule/variant is bit-exact with the oracle's
restatement of the Pannotia max rule."""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle
from conftest import csr_from_edges, load_fixtures, values_of

pytestmark = pytest.mark.gpu

FX = load_fixtures()
SCHEDS = [
    dict(schedule=gdb.SCHED_PERSISTENT, variant=gdb.VARIANT_PLAIN),
    dict(schedule=gdb.SCHED_PERSISTENT, variant=gdb.VARIANT_TUNED),
    dict(schedule=gdb.SCHED_RELAUNCH, variant=gdb.VARIANT_PLAIN),
    dict(schedule=gdb.SCHED_RELAUNCH, variant=gdb.VARIANT_PLAIN, relaunch_check=8),
]


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


def run(ctx, row, col, nv, **kw):
    g = gdb.DeviceCSR(gdb.CSR(row.size - 1, row, col, np.ones_like(col)), weights=False)
    vals = gdb.DeviceBuffer.from_numpy(nv.astype(np.float32))
    return gdb.color(ctx, g, vals, gdb.Opts.make(**kw))


@pytest.mark.parametrize("sched", SCHEDS, ids=lambda s: "-".join(map(str, s.values())))
@pytest.mark.parametrize("fx", FX["color"], ids=lambda f: f["name"])
def test_color_fixtures(ctx, fx, sched):
    row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    got, st = run(ctx, row, col, values_of(fx), wgs=64, **sched)
    assert got.tolist() == fx["color"]
    assert st.rounds == fx["rounds"]


@pytest.mark.parametrize("sched", SCHEDS, ids=lambda s: "-".join(map(str, s.values())))
@pytest.mark.parametrize("wgs", [64, 256, 1024])
def test_color_c1_rmat14(ctx, sched, wgs):
    g = gdb.gen_rmat(14, 16, seed=1)
    nv = gdb.pannotia_values(g.n)
    want, rounds = oracle.color_ref(g.row, g.col, nv)
    got, st = run(ctx, g.row, g.col, nv, wgs=wgs, **sched)
    print(sched, wgs, st.as_dict())
    assert (got == want).all()
    assert st.rounds == rounds


@pytest.mark.parametrize("bar", [gdb.BARRIER_XF, gdb.BARRIER_COUNTER, gdb.BARRIER_TREE, gdb.BARRIER_AUTO])
def test_color_permutation_values_all_barriers(ctx, bar):
    g = gdb.gen_rmat(13, 16, seed=1)
    nv = gdb.permutation(g.n, 1).astype(np.float32) / np.float32(g.n + 1)
    want, _ = oracle.color_ref(g.row, g.col, nv)
    ctx.configure(gdb.PROTOCOL_LOCKFREE, bar)
    try:
        for variant in (gdb.VARIANT_PLAIN, gdb.VARIANT_TUNED):
            got, _ = run(ctx, g.row, g.col, nv, wgs=256, variant=variant)
            assert (got == want).all()
    finally:
        ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)


def test_color_empty_graph(ctx):
    row = np.zeros(6, np.int32)
    got, st = run(ctx, row, np.zeros(0, np.int32), gdb.pannotia_values(5), wgs=64)
    assert got.tolist() == [1] * 5


def test_color_ties_report_error_not_hang(ctx):
    row, col, _ = csr_from_edges(3, [(0, 1), (1, 2), (0, 2)])
    with pytest.raises(gdb.GdbError):
        run(ctx, row, col, np.full(3, 0.5, np.float32), wgs=64)


def test_color_single_xcd(ctx):
    g = gdb.gen_rmat(14, 16, seed=1)
    nv = gdb.pannotia_values(g.n)
    want, rounds = oracle.color_ref(g.row, g.col, nv)
    ctx.set_xcd_only(0)
    try:
        for variant in (gdb.VARIANT_PLAIN, gdb.VARIANT_TUNED):
            got, st = run(ctx, g.row, g.col, nv, wgs=1024, groups=64, variant=variant)
            print("single-xcd", variant, st.as_dict())
            assert (got == want).all() and st.rounds == rounds
    finally:
        ctx.set_xcd_only(-1)


@pytest.mark.parametrize("order", ["index", "reversed"])
@pytest.mark.parametrize("shape", ["grid", "grid_holes", "rmat"])
def test_color_tuned_blockers(ctx, shape, order):
    # The tuned kernel re-walks a light row only when the neighbour holding
    # its maximum is coloured.  Index-ordered values make a grid colour as a
    # wavefront (a vertex waits on one neighbour for many rounds); R-MAT mixes
    # light rows with chunked heavy rows that can be blockers.
    g = {"grid": lambda: gdb.gen_grid(48, 48, 0.0, seed=1),
         "grid_holes": lambda: gdb.gen_grid(31, 57, 0.15, seed=2),
         "rmat": lambda: gdb.gen_rmat(11, 8, seed=5)}[shape]()
    nv = gdb.pannotia_values(g.n)
    if order == "reversed":
        nv = nv[::-1].copy()
    want, rr = oracle.color_ref(g.row, g.col, nv)
    got, st = run(ctx, np.asarray(g.row), np.asarray(g.col), nv, wgs=128, groups=16,
                  schedule=gdb.SCHED_PERSISTENT, variant=gdb.VARIANT_TUNED)
    assert (got == want).all()
    assert st.rounds == rr


# ---------------------------------------------------------------- max-min colouring
@pytest.mark.parametrize("wgs,groups", [(64, 3), (256, 0), (1024, 16)])
@pytest.mark.parametrize("graph", ["rmat14", "grid", "gnp", "hubs"])
def test_color_maxmin_matches_oracle(ctx, graph, wgs, groups):
    if graph == "rmat14":
        g = gdb.gen_rmat(14, 16, seed=1)  # the C1 graph
    elif graph == "grid":
        g = gdb.gen_grid(120, 90, 0.05, seed=2)
    elif graph == "gnp":
        g = gdb.gen_gnp(20000, 8.0, seed=3)
    else:  # R-MAT with heavy hub rows walked by whole waves
        g = gdb.gen_rmat(12, 64, seed=4)
    nv = (gdb.permutation(g.n, 9).astype(np.float64) / (g.n + 1)).astype(np.float32)
    want, rounds = oracle.color_maxmin_ref(g.row, g.col, nv)
    dg = gdb.DeviceCSR(g, weights=False)
    got, st = gdb.color_maxmin(ctx, dg, gdb.DeviceBuffer.from_numpy(nv), gdb.Opts.make(wgs=wgs, groups=groups))
    print(graph, wgs, groups, st.as_dict())
    assert (got == want).all()
    assert st.rounds == rounds


def test_color_maxmin_fewer_rounds_than_max_rule(ctx):
    g = gdb.gen_rmat(14, 16, seed=1)
    nv = (gdb.permutation(g.n, 1).astype(np.float64) / (g.n + 1)).astype(np.float32)
    dg = gdb.DeviceCSR(g, weights=False)
    d_nv = gdb.DeviceBuffer.from_numpy(nv)
    _, s_mm = gdb.color_maxmin(ctx, dg, d_nv, gdb.Opts.make(wgs=256, groups=64))
    _, s_max = gdb.color(ctx, dg, d_nv, gdb.Opts.make(wgs=256, groups=64, variant=gdb.VARIANT_TUNED))
    print("max-min rounds", s_mm.rounds, "max rounds", s_max.rounds)
    assert s_mm.rounds < s_max.rounds


def test_color_maxmin_ties_report_timeout(ctx):
    row, col, _ = csr_from_edges(4, [(0, 1), (1, 2), (2, 3)])
    dg = gdb.DeviceCSR(gdb.CSR(4, row, col, np.ones_like(col)), weights=False)
    with pytest.raises(gdb.GdbError) as e:
        gdb.color_maxmin(ctx, dg, gdb.DeviceBuffer.from_numpy(np.full(4, 0.5, np.float32)), gdb.Opts.make(wgs=64))
    assert e.value.code == gdb.GDB_E_TIMEOUT
