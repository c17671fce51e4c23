"""Oracle parity at the configurations' full sizes and at the launch shapes
the benchmarks run.

  grid 2048^2, 5% of edges removed (seed 2), weights U[1,255]: the batched
      engine at the benchmark's two tuned shapes (shared: wgs 128, 256 groups
      in two teams of 128, near/far band 4096, 256 local hops; owned: two
      teams of 128 vertex ranges, one group of 1024 threads per CU, band
      6144), the single-task tuned kernels, and the owner-partitioned solve
      at other shapes (256 ranges, or two teams of 128); BFS levels against
      the Lonestar worklist restatement, distances against Dijkstra, bit for
      bit
  the same shape on an R-MAT graph whose frontiers overflow the 2048-entry
      LDS queue (the spill path)
  R-MAT scale 22 (seed 3): MIS against greedy sequential MIS (bit for bit);
      BC from 16 sources against FP64 Brandes, |got - want| <= 1e-4 |want|
      per entry (floor 1e-6)
  4096 graphs (2048 G(n,p) + 2048 grids) x BFS, SSSP, colour in one launch
      at the batch benchmark's shape; every 16th graph against the oracle
  the learned selector (committed weights) on unseen R-MAT 20 and a 1024^2
      grid: BFS levels and SSSP distances bit for bit

Each test prints what it runs before the GPU call."""
import sys
import time
from pathlib import Path

import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
from bench import tuned_c2  # noqa: E402

# The benchmark's two engine shapes from the tuning table: the shared engine
# and the owner-partitioned one (the headline runs the faster, which the
# autotuner recorded).
TUNED_C2, _, _, _ = tuned_c2("shared")
OWNED_C2, OWNED_TEAM, OWNED_QCAP, _ = tuned_c2("owned")

BC_RTOL, BC_ATOL = 1e-4, 1e-6


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def grid2048():
    g = gdb.gen_grid(2048, 2048, 0.05, seed=2, wlo=1, whi=255)
    want_b, _ = oracle.bfs_worklist_ref(g.row, g.col, 0)
    want_s = oracle.dijkstra_tb(g.row, g.col, g.w, 0)
    return g, gdb.DeviceCSR(g), want_b, want_s


@pytest.fixture(scope="module")
def rmat22():
    t0 = time.time()
    g = gdb.gen_rmat(22, 16, seed=3)
    say(f"R-MAT 22: n={g.n} m={g.m} ({time.time() - t0:.1f}s)")
    return g, gdb.DeviceCSR(g, weights=False)


def batched_grid_solve(ctx, dg, shape, out_b, out_s):
    wgs, groups, delta, hops, mode = shape
    assert mode == gdb.SEL_NONE
    tasks = [gdb.make_task(gdb.TASK_BFS, dg, out_b, 0, hops=hops, transpose=dg),
             gdb.make_task(gdb.TASK_SSSP, dg, out_s, 0, delta=delta, hops=hops, transpose=dg)]
    return gdb.batch_solve(ctx, tasks, max(1, groups // 2), gdb.Opts.make(wgs=wgs, groups=groups))


def test_grid2048_batched_at_tuned_shape(ctx, grid2048):
    g, dg, want_b, want_s = grid2048
    out_b = gdb.DeviceBuffer.empty(g.n, np.uint32)
    out_s = gdb.DeviceBuffer.empty(g.n, np.uint32)
    for rep in range(2):
        say(f"grid 2048 batched, shape {TUNED_C2}, repeat {rep}")
        rounds, st = batched_grid_solve(ctx, dg, TUNED_C2, out_b, out_s)
        assert (out_b.to_numpy() == want_b).all()
        assert (out_s.to_numpy().astype(np.int64) == want_s).all()
        say(rounds, st.as_dict())
    assert st.participating == TUNED_C2[1]


def test_grid2048_owned_at_bench_shape(ctx, grid2048):
    g, dg, want_b, want_s = grid2048
    wgs, groups, delta, _, _ = OWNED_C2
    out_b = gdb.DeviceBuffer.empty(g.n, np.uint32)
    out_s = gdb.DeviceBuffer.empty(g.n, np.uint32)
    for rep in range(2):
        say(f"grid 2048 owned at the bench shape {OWNED_C2}, team {OWNED_TEAM}, repeat {rep}")
        tasks = [gdb.make_task(gdb.TASK_BFS, dg, out_b, 0, owned=True),
                 gdb.make_task(gdb.TASK_SSSP, dg, out_s, 0, delta=delta, owned=True)]
        _, st = gdb.batch_solve(ctx, tasks, OWNED_TEAM,
                                gdb.Opts.make(wgs=wgs, groups=groups, lds_queue=OWNED_QCAP))
        assert (out_b.to_numpy() == want_b).all()
        assert (out_s.to_numpy().astype(np.int64) == want_s).all()
        say(st.as_dict(), [t.work_items for t in tasks])


@pytest.mark.parametrize("team,wgs", [(256, 512), (128, 512)])
def test_grid2048_owned_ranges(ctx, grid2048, team, wgs):
    # Owner-partitioned solves at full size: 256 ranges of 16384 vertices
    # (row offsets in LDS) or two teams of 128 ranges of 32768 (without).
    g, dg, want_b, want_s = grid2048
    out_b = gdb.DeviceBuffer.empty(g.n, np.uint32)
    out_s = gdb.DeviceBuffer.empty(g.n, np.uint32)
    say(f"grid 2048 owned, team {team}, wgs {wgs}, 256 groups")
    tasks = [gdb.make_task(gdb.TASK_BFS, dg, out_b, 0, owned=True),
             gdb.make_task(gdb.TASK_SSSP, dg, out_s, 0, delta=4096, owned=True)]
    _, st = gdb.batch_solve(ctx, tasks, team, gdb.Opts.make(wgs=wgs, groups=256))
    assert (out_b.to_numpy() == want_b).all()
    assert (out_s.to_numpy().astype(np.int64) == want_s).all()
    say(st.as_dict(), [t.work_items for t in tasks])


@pytest.mark.parametrize("algo", ["bfs", "sssp"])
def test_grid2048_single_task_tuned(ctx, grid2048, algo):
    g, dg, want_b, want_s = grid2048
    o = gdb.Opts.make(variant=gdb.VARIANT_TUNED, wgs=512, groups=128)
    say(f"grid 2048 single-task tuned {algo}")
    if algo == "bfs":
        got, st = gdb.bfs_worklist(ctx, dg, 0, o)
        assert (got == want_b).all()
    else:
        got, st = gdb.sssp_worklist(ctx, dg, 0, o)
        assert (got.astype(np.int64) == want_s).all()
    say(st.as_dict())


def test_rmat_frontiers_spill_lds_queue_at_tuned_shape(ctx):
    # R-MAT 18 frontiers hold tens of thousands of vertices: at 256 local hops
    # a group's LDS queue (2048 entries) overflows and spills to the team queue.
    wgs, groups, delta, hops, _ = TUNED_C2
    g = gdb.gen_rmat(18, 16, seed=9, wlo=1, whi=255)
    dg = gdb.DeviceCSR(g)
    out_b = gdb.DeviceBuffer.empty(g.n, np.uint32)
    out_s = gdb.DeviceBuffer.empty(g.n, np.uint32)
    say(f"R-MAT 18 batched at shape {TUNED_C2}")
    rounds, st = batched_grid_solve(ctx, dg, TUNED_C2, out_b, out_s)
    assert (out_b.to_numpy() == oracle.bfs_tb(g.row, g.col, 0)).all()
    assert (out_s.to_numpy().astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all()
    say(rounds, st.as_dict())


@pytest.mark.parametrize("variant,wgs,groups", [(gdb.VARIANT_TUNED, 256, 2048), (gdb.VARIANT_PLAIN, 256, 0)])
def test_rmat22_mis(ctx, rmat22, variant, wgs, groups):
    g, dg = rmat22
    nv_h = (gdb.permutation(g.n, 3).astype(np.float64) / (g.n + 1)).astype(np.float32)
    say(f"R-MAT 22 MIS variant {variant} wgs {wgs} groups {groups}")
    got, st = gdb.mis(ctx, dg, gdb.DeviceBuffer.from_numpy(nv_h),
                      gdb.Opts.make(variant=variant, wgs=wgs, groups=groups))
    want = oracle.mis_tb(g.row, g.col, nv_h)
    assert (got == want).all()
    say(st.as_dict())


def test_rmat22_bc_16_sources(ctx, rmat22):
    g, dg = rmat22
    deg = np.diff(g.row)
    rng = np.random.default_rng(3)
    src = np.sort(rng.choice(np.flatnonzero(deg > 0), 16, replace=False)).astype(np.int32)
    say("R-MAT 22 BC, 16 sources, tuned (two eight-source sweeps)")
    got, st = gdb.bc(ctx, dg, dg, src, gdb.Opts.make(variant=gdb.VARIANT_TUNED, wgs=256, groups=2048))
    say(st.as_dict())
    t0 = time.time()
    want = oracle.bc_brandes_mt(g.row, g.col, src)
    say(f"oracle {time.time() - t0:.1f}s")
    err = np.abs(got.astype(np.float64) - want)
    ok = err <= np.maximum(BC_RTOL * np.abs(want), BC_ATOL)
    nz = want != 0
    say("max rel", float((err[nz] / want[nz]).max()), "nonzero", int(nz.sum()))
    assert ok.all(), (np.flatnonzero(~ok)[:10], got[~ok][:10], want[~ok][:10])


def test_c4_full_batch_sampled(ctx):
    rng = np.random.default_rng(4)
    t0 = time.time()
    gs = []
    for i in range(2048):
        n = int(np.floor(np.exp(rng.uniform(np.log(256), np.log(65536)))))
        gs.append(gdb.gen_gnp(n, 8.0, seed=4 + i, wlo=1, whi=100))
    for i in range(2048):
        sd = int(rng.integers(16, 257))
        gs.append(gdb.gen_grid(sd, sd, 0.0, seed=4 + 2048 + i, wlo=1, whi=100))
    say(f"4096 graphs generated ({time.time() - t0:.1f}s)")
    dgs = [gdb.DeviceCSR(g) for g in gs]
    nvs = [gdb.DeviceBuffer.from_numpy(gdb.pannotia_values(g.n)) for g in gs]
    outs = [[gdb.DeviceBuffer.empty(g.n, np.uint32), gdb.DeviceBuffer.empty(g.n, np.uint32),
             gdb.DeviceBuffer.empty(g.n, np.int32)] for g in gs]
    tasks = []
    for dg, nv, o in zip(dgs, nvs, outs):
        tasks.append(gdb.make_task(gdb.TASK_BFS, dg, o[0], 0))
        tasks.append(gdb.make_task(gdb.TASK_SSSP, dg, o[1], 0, delta=200))
        tasks.append(gdb.make_task(gdb.TASK_COLOR, dg, o[2], 0, node_value=nv))
    say("4096-graph batch: wgs 512, 1024 groups, teams of one group, band 200")
    rounds, st = gdb.batch_solve(ctx, tasks, 1, gdb.Opts.make(wgs=512, groups=1024))
    say(st.as_dict())
    for i in range(0, 4096, 16):
        g = gs[i]
        want_b, _ = oracle.bfs_worklist_ref(g.row, g.col, 0)
        assert (outs[i][0].to_numpy() == want_b).all(), ("bfs", i)
        assert (outs[i][1].to_numpy().astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all(), ("sssp", i)
        want_c, rc = oracle.color_ref(g.row, g.col, gdb.pannotia_values(g.n))
        assert (outs[i][2].to_numpy() == want_c).all(), ("color", i)
        assert rounds[3 * i + 2] == rc


@pytest.mark.parametrize("name", ["rmat20", "grid1024"])
def test_learned_selector_unseen_graphs(ctx, name):
    w = np.load(ROOT / "gpu_discovery_barrier_amd" / "selector_weights.npy").astype(np.uint16)
    weights = gdb.DeviceBuffer.from_numpy(w)
    if name == "rmat20":
        g = gdb.gen_rmat(20, 16, seed=50, wlo=1, whi=255)
    else:
        g = gdb.gen_grid(1024, 1024, 0.05, seed=7, wlo=1, whi=255)
    dg = gdb.DeviceCSR(g)
    sel = gdb.make_selector(gdb.SEL_MLP, narrow_groups=2, weights=weights)
    outs = [gdb.DeviceBuffer.empty(g.n, np.uint32) for _ in range(2)]
    tasks = [gdb.make_task(gdb.TASK_BFS, dg, outs[0], 0, transpose=dg),
             gdb.make_task(gdb.TASK_SSSP, dg, outs[1], 0, delta=512, transpose=dg)]
    say(f"learned selector on {name}: two teams of 128 groups x 256 threads")
    _, st = gdb.batch_solve(ctx, tasks, 128, gdb.Opts.make(wgs=256, groups=256), selector=sel)
    say(st.as_dict(), [list(t.arm_iters) for t in tasks])
    assert (outs[0].to_numpy() == oracle.bfs_tb(g.row, g.col, 0)).all()
    assert (outs[1].to_numpy().astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all()
    # The selector did choose: more than one arm ran on the R-MAT graph.
    if name == "rmat20":
        assert sum(1 for a in tasks[0].arm_iters if a > 0) >= 2
