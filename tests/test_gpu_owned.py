"""Owner-partitioned BFS / SSSP (gdb_task.owned): each group of a team keeps a
contiguous vertex range in LDS.  Levels and distances must equal the
oracle's (bfs_worklist_ref, dijkstra_tb) whatever the team size, group size,
queue tile, graph locality (grids: few cross-range edges; R-MAT: almost all)
or the ranges that stay empty (more groups than vertices)."""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


def solve(ctx, g, algos, team, wgs=256, groups=0, delta=0, source=0, qcap=0, owned=True):
    dg = gdb.DeviceCSR(g)
    outs, tasks = [], []
    for algo in algos:
        out = gdb.DeviceBuffer.empty(g.n, np.uint32)
        outs.append(out)
        tasks.append(gdb.make_task(algo, dg, out, source=source, delta=delta if algo == gdb.TASK_SSSP else 0,
                                   owned=owned))
    gdb.batch_solve(ctx, tasks, team, gdb.Opts.make(wgs=wgs, groups=groups or team, lds_queue=qcap))
    return [o.to_numpy() for o in outs], tasks


def check(g, algos, got, source=0):
    for algo, out in zip(algos, got):
        if algo == gdb.TASK_BFS:
            want, _ = oracle.bfs_worklist_ref(g.row, g.col, source)
            assert (out == want).all(), "bfs levels differ"
        else:
            want = oracle.dijkstra_tb(g.row, g.col, g.w, source)
            assert (out.astype(np.int64) == want).all(), "sssp distances differ"


BOTH = (gdb.TASK_BFS, gdb.TASK_SSSP)


@pytest.mark.parametrize("team,wgs", [(1, 256), (4, 64), (16, 256), (64, 1024)])
@pytest.mark.parametrize("delta", [0, 300])
def test_owned_grid_exact(ctx, team, wgs, delta):
    g = gdb.gen_grid(192, 160, 0.05, seed=11, wlo=1, whi=255)
    got, _ = solve(ctx, g, BOTH, team, wgs=wgs, delta=delta)
    check(g, BOTH, got)


@pytest.mark.parametrize("team", [8, 128])
def test_owned_rmat_cross_range_edges_exact(ctx, team):
    # Nearly every edge of an R-MAT graph crosses ranges: the output words
    # and the boundary scans carry the solve.
    g = gdb.gen_rmat(14, 8, seed=5, wlo=1, whi=100)
    got, _ = solve(ctx, g, BOTH, team, delta=200, source=3)
    check(g, BOTH, got, source=3)


def slot_edge_cases(n=3000, seed=21):
    # Out-degrees 0..7 (slot records with 0 to 4 edges, and the CSR walk from
    # 5 edges on), one vertex of degree 40 (walked by the whole wave), a
    # parallel edge in every record of 2+ edges (the vertex's first target
    # again, lighter, so a lane relaxes one vertex twice with two candidates),
    # a self-loop on every fifth vertex of degree 3+, and weights of 1.2e9 on
    # every 50th edge (at or above 1e9: they never improve a distance, in the
    # oracle as in the kernel).
    rng = np.random.default_rng(seed)
    deg = (np.arange(n) % 8).astype(np.int64)
    deg[17] = 40
    row = np.zeros(n + 1, np.int64)
    row[1:] = np.cumsum(deg)
    col = rng.integers(0, n, int(row[-1])).astype(np.int32)
    w = rng.integers(100, 201, col.size).astype(np.int32)
    for v in range(n):
        b, d = int(row[v]), int(deg[v])
        if d >= 2:
            col[b + 1] = col[b]
            w[b + 1] = rng.integers(1, 100)
        if d >= 3 and v % 5 == 0:
            col[b + 2] = v
    w[::50] = 1_200_000_000
    return gdb.CSR(n, row.astype(np.int32), col, w)


@pytest.mark.parametrize("team,wgs", [(1, 256), (4, 128), (16, 64)])
def test_owned_slot_record_edge_cases(ctx, team, wgs):
    g = slot_edge_cases()
    for source in (1, 17, 2999):
        got, _ = solve(ctx, g, BOTH, team, wgs=wgs, delta=150, source=source)
        check(g, BOTH, got, source=source)


def test_owned_queue_overflow_goes_through_far_set(ctx):
    # One group owns all of an R-MAT 13 graph with a 256-entry queue: the
    # hub's frontier overflows the queue into the far set.
    g = gdb.gen_rmat(13, 16, seed=9, wlo=1, whi=50)
    got, _ = solve(ctx, g, BOTH, 1, wgs=512, qcap=256)
    check(g, BOTH, got)


def test_owned_more_groups_than_vertices(ctx):
    g = gdb.gen_grid(5, 9, 0.0, seed=1, wlo=1, whi=9)
    got, _ = solve(ctx, g, BOTH, 64, wgs=64, source=7)
    check(g, BOTH, got, source=7)


def test_owned_disconnected_and_isolated_vertices(ctx):
    # Two grids side by side with no edge between them, plus the isolated
    # vertices a 30% removal leaves: unreached vertices stay at 1e9.
    g1 = gdb.gen_grid(40, 40, 0.3, seed=2, wlo=1, whi=30)
    n = 2 * g1.n
    row = np.concatenate([g1.row, g1.row[1:] + g1.m]).astype(np.int32)
    col = np.concatenate([g1.col, g1.col + g1.n]).astype(np.int32)
    w = np.concatenate([g1.w, g1.w]).astype(np.int32)
    g = gdb.CSR(n, row, col, w)
    got, _ = solve(ctx, g, BOTH, 8, delta=50)
    check(g, BOTH, got)
    assert (got[0][g1.n:] == 1000000000).all()


def test_owned_mixed_batch_with_shared_queue_tasks(ctx):
    # Owned and shared-queue tasks, and a colour task, in one launch.
    gs = [gdb.gen_grid(64, 64, 0.05, seed=3, wlo=1, whi=99), gdb.gen_gnp(3000, 6.0, seed=4, wlo=1, whi=99)]
    tasks, outs, keep = [], [], []
    for gi, g in enumerate(gs):
        dg = gdb.DeviceCSR(g)
        keep.append(dg)
        for algo, owned in ((gdb.TASK_BFS, True), (gdb.TASK_SSSP, gi == 0), (gdb.TASK_SSSP, False)):
            o = gdb.DeviceBuffer.empty(g.n, np.uint32)
            outs.append((gi, algo, o, None))
            tasks.append(gdb.make_task(algo, dg, o, delta=100 if algo == gdb.TASK_SSSP else 0, owned=owned))
        nv = gdb.DeviceBuffer.from_numpy(gdb.permutation(g.n, 7).astype(np.float32) / np.float32(g.n + 1))
        keep.append(nv)
        oc = gdb.DeviceBuffer.empty(g.n, np.int32)
        tasks.append(gdb.make_task(gdb.TASK_COLOR, dg, oc, node_value=nv))
        outs.append((gi, gdb.TASK_COLOR, oc, nv))
    rounds, _ = gdb.batch_solve(ctx, tasks, 4, gdb.Opts.make(wgs=256, groups=16))
    for (gi, algo, o, nv), r in zip(outs, rounds):
        g = gs[gi]
        if algo == gdb.TASK_COLOR:
            want, rr = oracle.color_ref(g.row, g.col, nv.to_numpy())
            assert (o.to_numpy() == want).all() and r == rr
        else:
            check(g, (algo,), [o.to_numpy()])


def test_owned_with_selector_runs_shared_queue_engine(ctx):
    g = gdb.gen_grid(48, 48, 0.05, seed=6, wlo=1, whi=40)
    dg = gdb.DeviceCSR(g)
    outs = [gdb.DeviceBuffer.empty(g.n, np.uint32) for _ in range(2)]
    tasks = [gdb.make_task(a, dg, o, delta=80 if a == gdb.TASK_SSSP else 0, owned=True, transpose=dg)
             for a, o in zip(BOTH, outs)]
    gdb.batch_solve(ctx, tasks, 4, gdb.Opts.make(wgs=256, groups=8), selector=gdb.make_selector(gdb.SEL_HEURISTIC, narrow_groups=2))
    check(g, BOTH, [o.to_numpy() for o in outs])


def test_owned_repeated_solves_reuse_state(ctx):
    # Termination counters and boundary marks are reset per task: three
    # solves in a row (different sources) on one context stay exact.
    g = gdb.gen_grid(100, 90, 0.05, seed=8, wlo=1, whi=200)
    for src in (0, 4321, 8999):
        got, _ = solve(ctx, g, BOTH, 32, delta=500, source=src)
        check(g, BOTH, got, source=src)


def test_owned_batch_graph_replays_match_oracle(ctx):
    # Owned BFS and SSSP captured as a HIP graph and replayed several times,
    # with the outputs scrambled and eager work on the same context between
    # replays: every replay recomputes exact levels and distances (the
    # solve's ranges, marks and epoch counters are set up inside the kernel).
    g = gdb.gen_grid(192, 160, 0.05, seed=12, wlo=1, whi=200)
    dg = gdb.DeviceCSR(g)
    outs = [gdb.DeviceBuffer.empty(g.n, np.uint32) for _ in BOTH]
    tasks = [gdb.make_task(algo, dg, out, 0, delta=400 if algo == gdb.TASK_SSSP else 0, owned=True)
             for algo, out in zip(BOTH, outs)]
    graph = gdb.BatchGraph(ctx, tasks, 16, gdb.Opts.make(wgs=256, groups=32, lds_queue=1024))
    try:
        for rep in range(3):
            for o in outs:
                junk = np.full(o.count, 7, dtype=o.dtype)
                gdb.check(gdb.lib().gdb_memcpy_htod(o.ptr, junk.ctypes.data_as(gdb.api.P), junk.nbytes), "htod")
            graph.launch()
            check(g, BOTH, [o.to_numpy() for o in outs])
            small = gdb.gen_grid(40, 40, 0.0, seed=rep)
            got, _ = solve(ctx, small, BOTH, 4, delta=100)  # eager work between replays
            check(small, BOTH, got)
    finally:
        graph.close()
