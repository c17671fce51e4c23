"""Batched engine: many tasks in one persistent launch, bit-exact with the
oracle (BFS levels, SSSP distances, colours), for teams of one group
(workgroup barriers only) and multi-group teams (team sub-barriers), with
and without the near/far SSSP split."""
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


def small_graphs(count, seed):
    rng = np.random.default_rng(seed)
    gs = []
    for i in range(count):
        if i % 2 == 0:
            n = int(np.exp(rng.uniform(np.log(64), np.log(4096))))
            gs.append(gdb.gen_gnp(n, 8.0, seed=seed * 1000 + i, wlo=1, whi=100))
        else:
            s = int(rng.integers(8, 48))
            gs.append(gdb.gen_grid(s, s, 0.0, seed=seed * 1000 + i, wlo=1, whi=100))
    return gs


def run_batch(ctx, gs, team_groups, delta, wgs=256, groups=0, hops=0):
    dgs, outs, vals, tasks, kinds = [], [], [], [], []
    for gi, g in enumerate(gs):
        dg = gdb.DeviceCSR(g)
        dgs.append(dg)
        nv = gdb.DeviceBuffer.from_numpy(gdb.permutation(g.n, gi + 1).astype(np.float32) / np.float32(g.n + 1))
        vals.append(nv)
        for algo in (gdb.TASK_BFS, gdb.TASK_SSSP, gdb.TASK_COLOR):
            out = gdb.DeviceBuffer.empty(g.n, np.int32 if algo == gdb.TASK_COLOR else np.uint32)
            outs.append(out)
            tasks.append(gdb.make_task(algo, dg, out, source=0, node_value=nv,
                                       delta=delta if algo == gdb.TASK_SSSP else 0,
                                       hops=hops if algo != gdb.TASK_COLOR else 0))
            kinds.append((gi, algo))
    rounds, st = gdb.batch_solve(ctx, tasks, team_groups, gdb.Opts.make(wgs=wgs, groups=groups))
    return kinds, outs, vals, rounds, st


def check(gs, kinds, outs, vals, rounds, label_correcting=False):
    for (gi, algo), out, r in zip(kinds, outs, rounds):
        g = gs[gi]
        got = out.to_numpy()
        if algo == gdb.TASK_BFS:
            want, iters = oracle.bfs_worklist_ref(g.row, g.col, 0)
            assert (got == want).all(), ("bfs", gi)
            if not label_correcting:
                assert r == iters
        elif algo == gdb.TASK_SSSP:
            want = oracle.dijkstra_tb(g.row, g.col, g.w, 0)
            assert (got.astype(np.int64) == want).all(), ("sssp", gi)
        else:
            nv = vals[gi].to_numpy()
            want, rr = oracle.color_ref(g.row, g.col, nv)
            assert (got == want).all(), ("color", gi)
            assert r == rr


@pytest.mark.parametrize("delta", [0, 64, 300])
def test_batch_single_group_teams(ctx, delta):
    gs = small_graphs(24, seed=4)
    kinds, outs, vals, rounds, st = run_batch(ctx, gs, 1, delta)
    print(st.as_dict())
    check(gs, kinds, outs, vals, rounds)


@pytest.mark.parametrize("hops", [0, 3, 16])
@pytest.mark.parametrize("team", [2, 8, 64])
def test_batch_multi_group_teams(ctx, team, hops):
    gs = [gdb.gen_grid(200, 200, 0.05, seed=2, wlo=1, whi=255), gdb.gen_rmat(13, 16, seed=7, wlo=1, whi=255)]
    kinds, outs, vals, rounds, st = run_batch(ctx, gs, team, 500, wgs=256, groups=4 * team, hops=hops)
    print(team, hops, st.as_dict())
    check(gs, kinds, outs, vals, rounds, label_correcting=hops > 0)


@pytest.mark.parametrize("hops", [1, 8, 64])
def test_batch_local_hops_single_group(ctx, hops):
    gs = small_graphs(12, seed=5)
    kinds, outs, vals, rounds, st = run_batch(ctx, gs, 1, 200, hops=hops)
    check(gs, kinds, outs, vals, rounds, label_correcting=True)


@pytest.mark.parametrize("hops", [1024, 4096])
def test_batch_large_hop_budgets(ctx, hops):
    # Budgets far above a grid's diameter: one group may carry the whole
    # wavefront through its LDS queue; distances stay exact.
    gs = [gdb.gen_grid(160, 160, 0.05, seed=3, wlo=1, whi=255), gdb.gen_grid(64, 64, 0.0, seed=4, wlo=1, whi=9)]
    kinds, outs, vals, rounds, st = run_batch(ctx, gs, 8, 1024, wgs=256, groups=16, hops=hops)
    check(gs, kinds, outs, vals, rounds, label_correcting=True)


def test_batch_rejects_hop_budget_above_limit(ctx):
    gs = small_graphs(1, seed=6)
    with pytest.raises(gdb.GdbError) as e:
        run_batch(ctx, gs, 1, 0, hops=4097)
    assert e.value.code == gdb.GDB_E_INVALID


def test_batch_local_hops_large_frontier_overflows_lds(ctx):
    # R-MAT frontiers exceed the 2048-entry LDS queue: the spill path runs.
    gs = [gdb.gen_rmat(15, 16, seed=11, wlo=1, whi=255)]
    kinds, outs, vals, rounds, st = run_batch(ctx, gs, 4, 0, wgs=1024, groups=8, hops=8)
    check(gs, kinds, outs, vals, rounds, label_correcting=True)


def test_batch_more_teams_than_tasks(ctx):
    gs = small_graphs(2, seed=9)
    kinds, outs, vals, rounds, st = run_batch(ctx, gs, 1, 100)
    check(gs, kinds, outs, vals, rounds)


def test_batch_matches_single_task_solvers(ctx):
    g = gdb.gen_grid(128, 128, 0.05, seed=2, wlo=1, whi=255)
    dg = gdb.DeviceCSR(g)
    a, _ = gdb.sssp_worklist(ctx, dg, 0, gdb.Opts.make(variant=gdb.VARIANT_TUNED))
    out = gdb.DeviceBuffer.empty(g.n, np.uint32)
    t = gdb.make_task(gdb.TASK_SSSP, dg, out, 0, delta=1000)
    gdb.batch_solve(ctx, [t], 16, gdb.Opts.make(wgs=512, groups=16))
    assert (out.to_numpy() == a).all()


def test_batch_graph_replays_match_eager(ctx):
    # The batch captured as a HIP graph and replayed several times, with
    # other (eager) work and synchronisations between replays: every replay
    # recomputes the outputs (they are scrambled first) and matches the
    # eager launch and the oracle.
    gs = small_graphs(10, seed=7)
    kinds, outs, vals, rounds, st = run_batch(ctx, gs, 2, 150, groups=16)
    eager = [o.to_numpy() for o in outs]
    check(gs, kinds, outs, vals, rounds)
    tasks, keep = [], []
    for (gi, algo), out in zip(kinds, outs):
        dg = gdb.DeviceCSR(gs[gi])
        keep.append(dg)
        tasks.append(gdb.make_task(algo, dg, out, 0, node_value=vals[gi],
                                   delta=150 if algo == gdb.TASK_SSSP else 0))
    graph = gdb.BatchGraph(ctx, tasks, 2, gdb.Opts.make(wgs=256, groups=16))
    try:
        for rep in range(4):
            for o in outs:
                o_np = np.full(o.count, 7, dtype=o.dtype)
                gdb.check(gdb.lib().gdb_memcpy_htod(o.ptr, o_np.ctypes.data_as(gdb.api.P), o_np.nbytes), "htod")
            r2, st2 = graph.launch()
            for o, e in zip(outs, eager):
                assert (o.to_numpy() == e).all(), rep
            assert r2 == rounds
            check(gs, kinds, outs, vals, r2)
            # Eager work on the same context between replays.
            run_batch(ctx, small_graphs(2, seed=100 + rep), 1, 0)
    finally:
        graph.close()



def hub_ring(n_ring, dup):
    # Vertex 0 links to every ring vertex (a row longer than one lane walks);
    # the ring vertices link to their two ring neighbours; `dup` repeats
    # every edge.  Symmetric.
    n = 1 + n_ring
    src, dst = [], []
    for i in range(1, n):
        j = 1 + (i % n_ring)
        src += [0, i, i, j]
        dst += [i, 0, j, i]
    src = np.repeat(np.array(src, np.int32), dup)
    dst = np.repeat(np.array(dst, np.int32), dup)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    row = np.zeros(n + 1, np.int32)
    np.add.at(row, src + 1, 1)
    return gdb.CSR(n, np.cumsum(row).astype(np.int32), dst, np.ones(dst.size, np.int32))


@pytest.mark.parametrize("team", [1, 4])
@pytest.mark.parametrize("order", ["index", "reversed", "random"])
def test_batch_color_blockers(ctx, team, order):
    # A vertex re-walks its row only when the neighbour holding its maximum
    # is coloured.  Index-ordered values make grids colour as a wavefront
    # (every vertex waits on one neighbour for many rounds); reversed and
    # random orders, duplicate edges and a hub row (reduced by the whole
    # wave) move the maxima around.
    gs = [gdb.gen_grid(40, 40, 0.0, seed=1), gdb.gen_grid(23, 61, 0.1, seed=2), hub_ring(100, 1), hub_ring(37, 2),
          gdb.gen_gnp(3000, 8.0, seed=3)]
    tasks, outs = [], []
    for gi, g in enumerate(gs):
        if order == "index":
            v = gdb.pannotia_values(g.n)
        elif order == "reversed":
            v = gdb.pannotia_values(g.n)[::-1].copy()
        else:
            v = gdb.permutation(g.n, 11 + gi).astype(np.float32) / np.float32(g.n + 1)
        nv = gdb.DeviceBuffer.from_numpy(v)
        out = gdb.DeviceBuffer.empty(g.n, np.int32)
        dg = gdb.DeviceCSR(g)
        tasks.append(gdb.make_task(gdb.TASK_COLOR, dg, out, source=0, node_value=nv))
        outs.append((g, v, out, nv, dg))
    rounds, _ = gdb.batch_solve(ctx, tasks, team, gdb.Opts.make(wgs=256, groups=8))
    for (g, v, out, _, _), r in zip(outs, rounds):
        want, rr = oracle.color_ref(g.row, g.col, v)
        assert (out.to_numpy() == want).all()
        assert r == rr


def test_batch_shape_and_lds_tile_from_tuning_table(ctx, tmp_path):
    # opts.wgs = 0: the launch shape, team size and LDS queue tile come from
    # the tuning table; a 256-int tile makes local hops spill early.
    from gpu_discovery_barrier_amd import autotune

    t = tmp_path / "tune.txt"
    autotune.write_table([autotune.row("batch", 1 << 20, 128, 8, 4, 0, 0, 256)], t)
    gs = [gdb.gen_grid(200, 200, 0.05, seed=2, wlo=1, whi=255), gdb.gen_rmat(12, 16, seed=7, wlo=1, whi=255)]
    try:
        gdb.tune_load(str(t))
        kinds, outs, vals, rounds, st = run_batch(ctx, gs, 0, 500, wgs=0, groups=0, hops=64)
    finally:
        gdb.tune_load(None)
    print(st.as_dict())
    assert 1 <= st.participating <= 8
    check(gs, kinds, outs, vals, rounds, label_correcting=True)


@pytest.mark.parametrize("qcap", [256, 1024, 4096])
def test_batch_lds_queue_tiles_are_exact(ctx, qcap):
    gs = [gdb.gen_grid(160, 160, 0.05, seed=3, wlo=1, whi=255), gdb.gen_rmat(14, 16, seed=11, wlo=1, whi=255)]
    dgs, outs, tasks = [], [], []
    for g in gs:
        dg = gdb.DeviceCSR(g)
        dgs.append(dg)
        ob, os_ = gdb.DeviceBuffer.empty(g.n, np.uint32), gdb.DeviceBuffer.empty(g.n, np.uint32)
        outs.append((ob, os_))
        tasks += [gdb.make_task(gdb.TASK_BFS, dg, ob, 0, hops=256), gdb.make_task(gdb.TASK_SSSP, dg, os_, 0, delta=1024,
                                                                                  hops=256)]
    gdb.batch_solve(ctx, tasks, 8, gdb.Opts.make(wgs=128, groups=32, lds_queue=qcap))
    for g, (ob, os_) in zip(gs, outs):
        assert (ob.to_numpy() == oracle.bfs_tb(g.row, g.col, 0)).all()
        assert (os_.to_numpy().astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all()


@pytest.mark.parametrize("qcap", [100, 8192])
def test_batch_rejects_lds_queue_outside_range(ctx, qcap):
    g = gdb.gen_grid(8, 8, 0.0, seed=1, wlo=1, whi=9)
    dg = gdb.DeviceCSR(g)
    out = gdb.DeviceBuffer.empty(g.n, np.uint32)
    with pytest.raises(gdb.GdbError) as e:
        gdb.batch_solve(ctx, [gdb.make_task(gdb.TASK_BFS, dg, out, 0)], 1, gdb.Opts.make(wgs=128, groups=2,
                                                                                      lds_queue=qcap))
    assert e.value.code == gdb.GDB_E_INVALID
