"""Degenerate inputs on every schedule: graphs without edges, isolated
vertices, a source with no out-edges, no BC sources.  Each result is checked
against the oracle (BC also against its closed form: zero everywhere)."""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle

pytestmark = pytest.mark.gpu

SCHEDS = [
    dict(schedule=gdb.SCHED_PERSISTENT, variant=gdb.VARIANT_PLAIN),
    dict(schedule=gdb.SCHED_PERSISTENT, variant=gdb.VARIANT_TUNED),
    dict(schedule=gdb.SCHED_RELAUNCH, variant=gdb.VARIANT_PLAIN),
]
SID = lambda s: "-".join(map(str, s.values()))


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


def edgeless(n):
    return gdb.CSR(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))


def star_with_isolated(n_leaves, n_isolated):
    # Vertex 0 links to the leaves (both directions); the last vertices have
    # no edges; vertex n-1 (isolated) is also used as a source.
    n = 1 + n_leaves + n_isolated
    src = np.concatenate([np.zeros(n_leaves, np.int32), np.arange(1, n_leaves + 1, dtype=np.int32)])
    dst = np.concatenate([np.arange(1, n_leaves + 1, dtype=np.int32), np.zeros(n_leaves, np.int32)])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    row = np.zeros(n + 1, np.int32)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    w = (np.arange(dst.size, dtype=np.int32) % 7) + 1
    return gdb.CSR(n, row, dst, w)


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("shape", ["edgeless", "star"])
def test_bfs_sssp_degenerate(ctx, sched, shape):
    g = edgeless(9) if shape == "edgeless" else star_with_isolated(20, 5)
    d = gdb.DeviceCSR(g)
    for src in (0, g.n - 1):  # a hub (or an edgeless vertex) and an isolated vertex
        got_b, _ = gdb.bfs_worklist(ctx, d, src, gdb.Opts.make(wgs=64, **sched))
        assert (got_b == oracle.bfs_tb(g.row, g.col, src)).all()
        got_s, _ = gdb.sssp_worklist(ctx, d, src, gdb.Opts.make(wgs=64, **sched))
        assert (got_s.astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, src)).all()


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("shape", ["edgeless", "star"])
def test_mis_degenerate(ctx, sched, shape):
    g = edgeless(9) if shape == "edgeless" else star_with_isolated(20, 5)
    nv = gdb.permutation(g.n, 5).astype(np.float32) / np.float32(g.n + 1)
    want, _ = oracle.mis_ref(g.row, g.col, nv)
    got, _ = gdb.mis(ctx, gdb.DeviceCSR(g), gdb.DeviceBuffer.from_numpy(nv), gdb.Opts.make(wgs=64, **sched))
    assert (got == want).all()
    if shape == "edgeless":
        assert (got == 2).all()  # every vertex joins


@pytest.mark.parametrize("sched", SCHEDS, ids=SID)
@pytest.mark.parametrize("shape", ["edgeless", "star"])
def test_bc_degenerate(ctx, sched, shape):
    g = edgeless(9) if shape == "edgeless" else star_with_isolated(20, 5)
    d = gdb.DeviceCSR(g)
    for sources in (np.zeros(0, np.int32), np.array([0, g.n - 1, 1], np.int32)):
        got, _ = gdb.bc(ctx, d, d, sources, gdb.Opts.make(wgs=64, **sched))
        want = np.asarray(oracle.bc_brandes_tb(g.row, g.col, sources), np.float64) if len(sources) else \
            np.zeros(g.n)
        assert np.allclose(got, want, rtol=1e-4, atol=1e-4 * max(1.0, float(np.abs(want).max())))
        if shape == "edgeless" or len(sources) == 0:
            assert (got == 0).all()


@pytest.mark.parametrize("hops", [0, 8])
def test_batch_degenerate_tasks(ctx, hops):
    gs = [edgeless(9), star_with_isolated(20, 5), edgeless(1)]
    tasks, outs, vals = [], [], []
    dgs = [gdb.DeviceCSR(g) for g in gs]
    for g, dg in zip(gs, dgs):
        nv = gdb.DeviceBuffer.from_numpy(gdb.pannotia_values(g.n))
        vals.append(nv)
        for algo in (gdb.TASK_BFS, gdb.TASK_SSSP, gdb.TASK_COLOR):
            out = gdb.DeviceBuffer.empty(g.n, np.int32 if algo == gdb.TASK_COLOR else np.uint32)
            outs.append((g, algo, out, nv))
            tasks.append(gdb.make_task(algo, dg, out, source=g.n - 1, node_value=nv,
                                       delta=5 if algo == gdb.TASK_SSSP else 0,
                                       hops=hops if algo != gdb.TASK_COLOR else 0))
    gdb.batch_solve(ctx, tasks, 2, gdb.Opts.make(wgs=64, groups=4))
    for g, algo, out, nv in outs:
        got = out.to_numpy()
        if algo == gdb.TASK_BFS:
            assert (got == oracle.bfs_tb(g.row, g.col, g.n - 1)).all()
        elif algo == gdb.TASK_SSSP:
            assert (got.astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, g.n - 1)).all()
        else:
            want, _ = oracle.color_ref(g.row, g.col, nv.to_numpy())
            assert (got == want).all()
