"""Parity of MIS, pull SSSP, BC and the worklist BFS / SSSP with the oracle,
for every schedule, through the C ABI.

Bars: MIS sets, SSSP distances and BFS levels bit-exact; BC (FP32) within
1e-4 relative of the FP64 Brandes oracle entry by entry, |got - want| <=
1e-4 * |want|, with an absolute floor of 1e-6 that only matters for entries
that are (near) zero."""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle
from conftest import csr_from_edges, grid_csr, load_fixtures, values_of

pytestmark = pytest.mark.gpu
FX = load_fixtures()

SCHEDS = [
    dict(schedule=gdb.SCHED_PERSISTENT, variant=gdb.VARIANT_PLAIN),
    dict(schedule=gdb.SCHED_PERSISTENT, variant=gdb.VARIANT_TUNED),
    dict(schedule=gdb.SCHED_RELAUNCH, variant=gdb.VARIANT_PLAIN),
    dict(schedule=gdb.SCHED_RELAUNCH, variant=gdb.VARIANT_PLAIN, relaunch_check=5),
]
SID = lambda s: "-".join(map(str, s.values()))
BC_RTOL = 1e-4
BC_ATOL = 1e-6


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


def dev(row, col, w=None):
    w = np.ones_like(col) if w is None else w
    return gdb.DeviceCSR(gdb.CSR(row.size - 1, row, col, w))


def bc_close(got, want):
    """Per-entry relative bar: |got - want| <= 1e-4 |want|, floor 1e-6."""
    want = np.asarray(want, np.float64)
    err = np.abs(got.astype(np.float64) - want)
    ok = err <= np.maximum(BC_RTOL * np.abs(want), BC_ATOL)
    assert ok.all(), (np.flatnonzero(~ok)[:10], got[~ok][:10], want[~ok][:10])
    nz = want != 0
    return float((err[nz] / np.abs(want[nz])).max()) if nz.any() else 0.0


# ---------------------------------------------------------------- MIS
@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("fx", FX["mis"], ids=lambda f: f["name"])
def test_mis_fixtures(ctx, fx, sched):
    row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    got, _ = gdb.mis(ctx, dev(row, col), gdb.DeviceBuffer.from_numpy(values_of(fx)),
                     gdb.Opts.make(wgs=64, **sched))
    assert got.tolist() == fx["s"]


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("vals", ["id", "perm"])
def test_mis_rmat(ctx, sched, vals):
    g = gdb.gen_rmat(15, 16, seed=3)
    nv = gdb.pannotia_values(g.n) if vals == "id" else \
        gdb.permutation(g.n, 3).astype(np.float32) / np.float32(g.n + 1)
    want, rounds = oracle.mis_ref(g.row, g.col, nv)
    got, st = gdb.mis(ctx, dev(g.row, g.col), gdb.DeviceBuffer.from_numpy(nv), gdb.Opts.make(**sched))
    print(vals, SID(sched), st.as_dict())
    assert (got == want).all()
    assert st.rounds == rounds


# ---------------------------------------------------------------- pull SSSP
@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("fx", FX["sssp_pull"], ids=lambda f: f["name"])
def test_sssp_pull_fixtures(ctx, fx, sched):
    row, col, w = csr_from_edges(fx["n"], fx["edges"], fx["w"])
    got, _ = gdb.sssp_pull(ctx, dev(row, col, w), 0, gdb.Opts.make(wgs=64, **sched))
    assert got.tolist() == fx["dist"]


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
def test_sssp_pull_rmat(ctx, sched):
    g = gdb.gen_rmat(14, 16, seed=5, wlo=1, whi=100)
    gt = g.transpose()
    want, rounds = oracle.sssp_pull_ref(gt.row, gt.col, gt.w, 0)
    got, st = gdb.sssp_pull(ctx, dev(gt.row, gt.col, gt.w), 0, gdb.Opts.make(**sched))
    print(SID(sched), st.as_dict())
    assert (got == want).all()
    assert st.rounds == rounds


@pytest.mark.parametrize("sched", SCHEDS[:2], ids=SID)
def test_sssp_pull_round_cap_grid(ctx, sched):
    # 30x3 grid: hop diameter 31 > n-1?  No: n = 90; use a path so the cap binds.
    n = 40
    row, col, w = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)], [3] * (n - 1))
    want, rounds = oracle.sssp_pull_ref(row, col, w, 0)
    got, st = gdb.sssp_pull(ctx, dev(row, col, w), 0, gdb.Opts.make(wgs=64, **sched))
    assert (got == want).all() and want[-1] == oracle.PANNOTIA_INF and st.rounds == rounds


# ---------------------------------------------------------------- BC
@pytest.mark.parametrize("sched", SCHEDS[:3], ids=SID)
@pytest.mark.parametrize("fx", FX["bc"], ids=lambda f: f["name"])
def test_bc_fixtures(ctx, fx, sched):
    row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    g = dev(row, col)
    got, _ = gdb.bc(ctx, g, g, np.array(fx["sources"], np.int32), gdb.Opts.make(wgs=64, **sched))
    bc_close(got, fx["bc"])


@pytest.mark.parametrize("sched", SCHEDS[:3], ids=SID)
def test_bc_rmat(ctx, sched):
    g = gdb.gen_rmat(13, 16, seed=3)
    src = gdb.permutation(g.n, 3)[:16]
    want = oracle.bc_brandes_tb(g.row, g.col, src)
    d = dev(g.row, g.col)
    got, st = gdb.bc(ctx, d, d, src, gdb.Opts.make(**sched))
    print(SID(sched), st.as_dict())
    bc_close(got, want)


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
def test_bc_rmat15_64_sources(ctx, sched):
    # R-MAT 15 with 64 sources: eight full eight-source sweeps on the tuned
    # schedule, 64 single-source passes on the others.
    g = gdb.gen_rmat(15, 16, seed=3)
    deg = np.diff(g.row)
    src = gdb.permutation(g.n, 15)
    src = src[deg[src] > 0][:64].astype(np.int32)
    want = oracle.bc_brandes_tb(g.row, g.col, src)
    d = dev(g.row, g.col)
    got, st = gdb.bc(ctx, d, d, src, gdb.Opts.make(**sched))
    rel = bc_close(got, want)
    print(SID(sched), st.as_dict(), "max rel", rel)
    assert (want > 0).sum() > g.n // 4  # most vertices carry a nonzero centrality


def test_bc_close_rejects_small_entries_off():
    # The bar must catch a wrong small entry: zeroing the smallest nonzero
    # oracle value fails it, as does a 2e-4 relative error on one entry.
    g = gdb.gen_rmat(10, 16, seed=3)
    src = gdb.permutation(g.n, 3)[:8]
    want = oracle.bc_brandes_tb(g.row, g.col, src)
    k = int(np.argmin(np.where(want > 0, want, np.inf)))
    bad = want.astype(np.float32).copy()
    bad[k] = 0
    with pytest.raises(AssertionError):
        bc_close(bad, want)
    j = int(np.argmax(want))
    bad = want.astype(np.float32).copy()
    bad[j] *= 1 + 2e-4
    with pytest.raises(AssertionError):
        bc_close(bad, want)


def test_bc_sweeps_partial_batch_and_repeated_sources(ctx):
    # Eight sources per sweep: 13 sources leave a partial last sweep, and a
    # source listed twice (within one sweep and across two) is counted twice,
    # as the single-source kernels and the oracle count it.
    g = gdb.gen_rmat(12, 16, seed=21)
    p = gdb.permutation(g.n, 21)
    src = np.concatenate([p[:5], p[:1], p[5:11], p[1:2]]).astype(np.int32)  # repeats in and across sweeps
    d = dev(g.row, g.col)
    want = oracle.bc_brandes_tb(g.row, g.col, src)
    for sched in SCHEDS[:3]:
        got, _ = gdb.bc(ctx, d, d, src, gdb.Opts.make(**sched))
        bc_close(got, want)


def test_bc_tuned_is_deterministic(ctx):
    g = gdb.gen_rmat(12, 16, seed=8)
    src = gdb.permutation(g.n, 8)[:8]
    d = dev(g.row, g.col)
    a, _ = gdb.bc(ctx, d, d, src, gdb.Opts.make(variant=gdb.VARIANT_TUNED))
    b, _ = gdb.bc(ctx, d, d, src, gdb.Opts.make(variant=gdb.VARIANT_TUNED, wgs=1024, groups=7))
    assert (a == b).all()


# ---------------------------------------------------------------- worklist BFS / SSSP
@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("fx", FX["bfs"], ids=lambda f: f["name"])
def test_bfs_fixtures(ctx, fx, sched):
    if "rows" in fx:
        row, col, _ = grid_csr(fx["rows"], fx["cols"])
    else:
        row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    got, _ = gdb.bfs_worklist(ctx, dev(row, col), 0, gdb.Opts.make(wgs=64, **sched))
    assert got.tolist() == fx["dist"]


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("fx", FX["sssp_worklist"], ids=lambda f: f["name"])
def test_sssp_worklist_fixtures(ctx, fx, sched):
    row, col, w = csr_from_edges(fx["n"], fx["edges"], fx["w"])
    got, _ = gdb.sssp_worklist(ctx, dev(row, col, w), 0, gdb.Opts.make(wgs=64, **sched))
    assert got.tolist() == fx["dist"]


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
def test_bfs_sssp_grid(ctx, sched):
    g = gdb.gen_grid(256, 256, 0.05, seed=2, wlo=1, whi=255)
    d = dev(g.row, g.col, g.w)
    want_b, iters = oracle.bfs_worklist_ref(g.row, g.col, 0)
    got_b, st = gdb.bfs_worklist(ctx, d, 0, gdb.Opts.make(**sched))
    print("bfs", SID(sched), st.as_dict())
    assert (got_b == want_b).all()
    assert st.rounds == iters
    want_s, _ = oracle.sssp_worklist_ref(g.row, g.col, g.w, 0)
    got_s, st = gdb.sssp_worklist(ctx, d, 0, gdb.Opts.make(**sched))
    print("sssp", SID(sched), st.as_dict())
    assert (got_s == want_s).all()


@pytest.mark.parametrize("sched", SCHEDS[:2], ids=SID)
def test_bfs_sssp_rmat(ctx, sched):
    g = gdb.gen_rmat(15, 16, seed=6, wlo=1, whi=255)
    d = dev(g.row, g.col, g.w)
    got_b, _ = gdb.bfs_worklist(ctx, d, 0, gdb.Opts.make(**sched))
    assert (got_b == oracle.bfs_tb(g.row, g.col, 0)).all()
    got_s, _ = gdb.sssp_worklist(ctx, d, 0, gdb.Opts.make(**sched))
    assert (got_s.astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all()


def test_verify_distances_kernel(ctx):
    g = gdb.gen_grid(64, 64, 0.05, seed=2, wlo=1, whi=255)
    d = dev(g.row, g.col, g.w)
    want, _ = oracle.sssp_worklist_ref(g.row, g.col, g.w, 0)
    good = gdb.DeviceBuffer.from_numpy(want)
    assert gdb.verify_distances(ctx, d, good, weighted=True) == 0
    bad = want.copy()
    bad[100] += 1000
    n_bad = oracle.verify_distances_ref(g.row, g.col, g.w, bad)
    assert n_bad > 0
    assert gdb.verify_distances(ctx, d, gdb.DeviceBuffer.from_numpy(bad), weighted=True) == n_bad


def test_bc_directed_graph_uses_transpose(ctx):
    # Arcs u -> v with u < v only: the in-edge CSR differs from the CSR, so
    # the tuned kernel takes its two-walk path (pull in-edges, push out-edges).
    g = gdb.gen_rmat(11, 16, seed=12)
    src = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.row))
    keep = src < g.col
    row = np.zeros(g.n + 1, np.int32)
    np.add.at(row, src[keep] + 1, 1)
    dg_h = gdb.CSR(g.n, np.cumsum(row).astype(np.int32), g.col[keep].astype(np.int32),
                   g.w[keep].astype(np.int32))
    gt_h = dg_h.transpose()
    sources = np.array([0, 1, 2, 5, 9, 17], np.int32)
    want = oracle.bc_brandes_tb(dg_h.row, dg_h.col, sources)
    d, dt = dev(dg_h.row, dg_h.col), dev(gt_h.row, gt_h.col)
    for sched in SCHEDS[:3]:
        got, _ = gdb.bc(ctx, d, dt, sources, gdb.Opts.make(**sched))
        bc_close(got, want)


def test_bc_symmetric_walk_matches_two_walks(ctx):
    # The same symmetric graph passed once (one walk per frontier vertex) and
    # as a separate transpose copy (two walks) gives the same centralities.
    g = gdb.gen_rmat(12, 16, seed=13)
    src = gdb.permutation(g.n, 13)[:8]
    d1, d2 = dev(g.row, g.col), dev(g.row, g.col)
    o = gdb.Opts.make(variant=gdb.VARIANT_TUNED)
    a, _ = gdb.bc(ctx, d1, d1, src, o)
    b, _ = gdb.bc(ctx, d1, d2, src, o)
    bc_close(a, oracle.bc_brandes_tb(g.row, g.col, src))
    bc_close(b, oracle.bc_brandes_tb(g.row, g.col, src))


def symmetric_multigraph(scale, seed):
    # An R-MAT graph with every fifth undirected edge doubled (both arcs) and
    # a self-loop on every seventh vertex; still symmetric.
    g = gdb.gen_rmat(scale, 16, seed=seed)
    src = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.row))
    dst = np.asarray(g.col, np.int32)
    fwd = np.flatnonzero(src < dst)[::5]
    loops = np.arange(0, g.n, 7, dtype=np.int32)
    s = np.concatenate([src, src[fwd], dst[fwd], loops])
    d = np.concatenate([dst, dst[fwd], src[fwd], loops])
    order = np.lexsort((d, s))
    s, d = s[order], d[order]
    row = np.zeros(g.n + 1, np.int32)
    np.add.at(row, s + 1, 1)
    return np.cumsum(row).astype(np.int32), d.astype(np.int32)


def test_bc_multigraph_duplicate_edges_and_self_loops(ctx):
    # Parallel edges count once per edge in sigma and delta (as in the
    # oracle); a self-loop never joins a level.  Eleven sources: a full
    # eight-source sweep and a partial one, one source with a self-loop.
    row, col = symmetric_multigraph(11, 4)
    n = row.size - 1
    src = np.concatenate([[0, 7], gdb.permutation(n, 4)[:9]]).astype(np.int32)
    want = oracle.bc_brandes_tb(row, col, src)
    d = dev(row, col)
    for sched in SCHEDS[:3]:
        got, _ = gdb.bc(ctx, d, d, src, gdb.Opts.make(**sched))
        bc_close(got, want)


@pytest.mark.parametrize("sched", SCHEDS[:3], ids=SID)
def test_worklist_multigraph_parallel_edges_with_different_weights(ctx, sched):
    # Parallel edges carry different weights, so two slots of one lane can
    # relax the same vertex with different candidates in one step.
    row, col = symmetric_multigraph(11, 6)
    w = (np.arange(col.size, dtype=np.int64) * 7919 % 97 + 1).astype(np.int32)
    d = dev(row, col, w)
    for source in (0, 5, 1000):
        got_b, _ = gdb.bfs_worklist(ctx, d, source, gdb.Opts.make(**sched))
        assert (got_b == oracle.bfs_tb(row, col, source)).all()
        got_s, _ = gdb.sssp_worklist(ctx, d, source, gdb.Opts.make(**sched))
        assert (got_s.astype(np.int64) == oracle.dijkstra_tb(row, col, w, source)).all()


@pytest.mark.parametrize("hops", [0, 8])
def test_batch_multigraph_parallel_edges_with_different_weights(ctx, hops):
    row, col = symmetric_multigraph(11, 6)
    w = (np.arange(col.size, dtype=np.int64) * 7919 % 97 + 1).astype(np.int32)
    d = dev(row, col, w)
    ob = gdb.DeviceBuffer.empty(row.size - 1, np.uint32)
    os_ = gdb.DeviceBuffer.empty(row.size - 1, np.uint32)
    tasks = [gdb.make_task(gdb.TASK_BFS, d, ob, 3, hops=hops),
             gdb.make_task(gdb.TASK_SSSP, d, os_, 3, delta=40, hops=hops)]
    gdb.batch_solve(ctx, tasks, 2, gdb.Opts.make(wgs=128, groups=4))
    assert (ob.to_numpy() == oracle.bfs_tb(row, col, 3)).all()
    assert (os_.to_numpy().astype(np.int64) == oracle.dijkstra_tb(row, col, w, 3)).all()
