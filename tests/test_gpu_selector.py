"""Per-iteration schedule selector of the batched engine, and its MLP.

Parity: whatever arm each iteration runs (push / pull, whole team / narrow
subset of the team), BFS levels and SSSP distances equal the oracle's bit
for bit; the HIP forward and backward passes of the MLP match the FP64
oracle (tolerances written in the tests, fp32 accumulation vs FP64).
"""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle
from gpu_discovery_barrier_amd.selector_train import balanced, steady_records

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


def directed_half(g):
    """Keeps the arcs u -> v with u < v (a directed graph whose transpose differs)."""
    src = np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.row))
    keep = src < g.col
    s, d, w = src[keep], g.col[keep], g.w[keep]
    row = np.zeros(g.n + 1, np.int32)
    np.add.at(row, s + 1, 1)
    return gdb.CSR(g.n, np.cumsum(row).astype(np.int32), d.astype(np.int32), w.astype(np.int32))


def graphs():
    return {
        "rmat12": gdb.gen_rmat(12, 16, seed=3, wlo=1, whi=255),
        "grid96": gdb.gen_grid(96, 96, 0.05, seed=2, wlo=1, whi=255),
        "gnp3000": gdb.gen_gnp(3000, 8.0, seed=4, wlo=1, whi=100),
        "rmat11_directed": directed_half(gdb.gen_rmat(11, 16, seed=5, wlo=1, whi=255)),
    }


GRAPHS = graphs()


def run(ctx, g, sel, team=8, wgs=256, delta=0, algos=(gdb.TASK_BFS, gdb.TASK_SSSP)):
    dg = gdb.DeviceCSR(g)
    dt = gdb.DeviceCSR(g.transpose())
    outs, tasks = [], []
    for a in algos:
        out = gdb.DeviceBuffer.empty(g.n, np.uint32)
        outs.append(out)
        tasks.append(gdb.make_task(a, dg, out, 0, delta=delta if a == gdb.TASK_SSSP else 0, transpose=dt))
    rounds, st = gdb.batch_solve(ctx, tasks, team, gdb.Opts.make(wgs=wgs, groups=team * len(algos)), selector=sel)
    return [o.to_numpy() for o in outs], tasks, st


def check(g, algos, outs):
    for a, got in zip(algos, outs):
        if a == gdb.TASK_BFS:
            want, _ = oracle.bfs_worklist_ref(g.row, g.col, 0)
        else:
            want = oracle.dijkstra_tb(g.row, g.col, g.w, 0)
        assert (got.astype(np.int64) == want).all(), a


@pytest.mark.parametrize("arm", [gdb.ARM_PUSH, gdb.ARM_PULL, gdb.ARM_PUSH_NARROW, gdb.ARM_PULL_NARROW])
@pytest.mark.parametrize("name", list(GRAPHS))
def test_fixed_arm_parity(ctx, name, arm):
    g = GRAPHS[name]
    sel = gdb.make_selector(gdb.SEL_FIXED, fixed_arm=arm, narrow_groups=2)
    algos = (gdb.TASK_BFS, gdb.TASK_SSSP)
    outs, tasks, st = run(ctx, g, sel, team=8, delta=0 if "grid" in name else 300)
    print(name, arm, [list(t.arm_iters) for t in tasks], st.kernel_ms)
    check(g, algos, outs)
    for t in tasks:
        assert t.arm_iters[arm] == sum(t.arm_iters) > 0


@pytest.mark.parametrize("team,narrow", [(1, 1), (4, 1), (16, 3), (64, 1)])
@pytest.mark.parametrize("mode", [gdb.SEL_RANDOM, gdb.SEL_HEURISTIC])
def test_switching_arms_parity(ctx, mode, team, narrow):
    # Random arms redraw one iteration in eight: on the grid's ~190 levels
    # the schedule (and team width) switches dozens of times, so parked
    # groups are woken and parked again many times.
    for seed, name in enumerate(["rmat12", "grid96", "rmat11_directed"]):
        g = GRAPHS[name]
        sel = gdb.make_selector(mode, narrow_groups=narrow, seed=seed + 11)
        outs, tasks, st = run(ctx, g, sel, team=team, delta=500)
        check(g, (gdb.TASK_BFS, gdb.TASK_SSSP), outs)
        if mode == gdb.SEL_RANDOM and team > narrow and name == "grid96":
            assert sum(1 for a in range(4) if tasks[0].arm_iters[a] > 0) >= 3


def test_trace_records(ctx):
    g = GRAPHS["rmat12"]
    cap = 4096
    trace = gdb.DeviceBuffer.from_numpy(np.zeros(cap * gdb.TRACE_STRIDE, np.float32))
    cnt = gdb.DeviceBuffer.from_numpy(np.zeros(1, np.int32))
    sel = gdb.make_selector(gdb.SEL_RANDOM, narrow_groups=2, seed=3, trace=trace, trace_count=cnt)
    outs, tasks, st = run(ctx, g, sel, team=8)
    n = int(cnt.to_numpy()[0])
    rec = trace.to_numpy().reshape(cap, gdb.TRACE_STRIDE)[:n]
    iters = sum(sum(t.arm_iters) for t in tasks)
    assert 0 < n <= iters
    assert set(np.unique(rec[:, 16]).astype(int)) <= {0, 1, 2, 3}
    assert np.all((rec[:, 18] == 0) | (rec[:, 18] == 1))
    # Every record but each task's last has a measured cost.
    assert (rec[:, 17] > 0).sum() >= n - 2
    assert np.all(np.isfinite(rec[:, :16])) and np.all(rec[:, :16] >= -1) and np.all(rec[:, :16] <= 2)


def test_mlp_forward_matches_oracle(ctx):
    p = gdb.mlp_init(7)
    wb = gdb.bf16_round(p)
    x = np.random.default_rng(1).uniform(0, 1, (64, 16)).astype(np.float32)
    got = gdb.mlp_forward(ctx, gdb.DeviceBuffer.from_numpy(wb), x)
    want = oracle.mlp_forward(gdb.bf16_to_f32(wb).astype(np.float64), x.astype(np.float64))
    # fp32 accumulation of 16 + 64 + 64 products vs FP64.
    assert np.allclose(got, want, rtol=1e-4, atol=1e-4)


def synthetic_records(count, seed):
    rng = np.random.default_rng(seed)
    rec = np.zeros((count, gdb.TRACE_STRIDE), np.float32)
    rec[:, :16] = rng.uniform(0, 1, (count, 16))
    rec[:, 16] = rng.integers(0, 4, count)
    # A cost that depends on features and arm (so the network can learn it).
    arm = rec[:, 16].astype(int)
    rec[:, 17] = np.exp2(2 + 4 * rec[:, 0] + 3 * (arm == 1) * rec[:, 1] - 2 * (arm == 2) * rec[:, 2] + arm)
    rec[::9, 17] = -1
    return rec


def test_mlp_gradient_matches_oracle(ctx):
    p = gdb.mlp_init(8)
    wb = gdb.bf16_round(p)
    rec = synthetic_records(3000, 2)
    g, loss, used = gdb.mlp_grad(ctx, gdb.DeviceBuffer.from_numpy(wb), gdb.DeviceBuffer.from_numpy(rec), len(rec))
    gw, lw, uw = oracle.mlp_grad(gdb.bf16_to_f32(wb).astype(np.float64), rec)
    assert used == uw
    assert loss == pytest.approx(lw, rel=1e-4)
    # fp32 gradient sums (atomic order varies) vs FP64: relative to the norm.
    err = np.abs(g - gw).max() / np.abs(gw).max()
    print("grad rel err", err)
    assert err < 1e-4
    assert np.linalg.norm(g - gw) / np.linalg.norm(gw) < 1e-4


def test_mlp_training_reduces_loss(ctx):
    rec = synthetic_records(20000, 3)
    drec = gdb.DeviceBuffer.from_numpy(rec)
    p0 = gdb.mlp_init(9)
    _, l0, _ = gdb.mlp_grad(ctx, gdb.DeviceBuffer.from_numpy(gdb.bf16_round(p0)), drec, len(rec))
    p, w, l1 = gdb.mlp_train(ctx, p0, drec, len(rec), steps=1500, batch=256, lr=3e-3, seed=1)
    print("loss", l0, "->", l1)
    assert l1 < 0.1 * l0 and l1 < 0.5
    # The trained bf16 weights are the rounding of the fp32 master copy.
    assert (w.to_numpy() == gdb.bf16_round(p)).all()


def solve_one(ctx, gr, dg, algo, sel, team=64):
    out = gdb.DeviceBuffer.empty(gr.n, np.uint32)
    t = gdb.make_task(algo, dg, out, 0, delta=512 if algo == gdb.TASK_SSSP else 0, transpose=dg)
    best = None
    for _ in range(3):
        _, st = gdb.batch_solve(ctx, [t], team, gdb.Opts.make(wgs=256, groups=team), selector=sel)
        best = st.kernel_ms if best is None else min(best, st.kernel_ms)
    return out.to_numpy(), best, list(t.arm_iters)


def test_trained_selector_gains_over_plain_schedule(ctx):
    """End to end: traces logged by the engine (random arms) on training
    graphs -> Adam on the device -> the MLP picks arms on unseen graphs.
    Outputs stay bit-exact; the learned schedule is faster than the plain
    push schedule (the method without the selector) on unseen R-MAT graphs,
    where bottom-up iterations pay off."""
    cap = 200000
    trace = gdb.DeviceBuffer.from_numpy(np.zeros(cap * gdb.TRACE_STRIDE, np.float32))
    cnt = gdb.DeviceBuffer.from_numpy(np.zeros(1, np.int32))
    runs = 0
    bounds = [0]
    for g in (gdb.gen_rmat(13, 16, seed=21, wlo=1, whi=255), gdb.gen_rmat(15, 16, seed=22, wlo=1, whi=255),
              gdb.gen_grid(160, 160, 0.05, seed=23, wlo=1, whi=255)):
        dg = gdb.DeviceCSR(g)
        for rep in range(3):
            for algo in (gdb.TASK_BFS, gdb.TASK_SSSP):
                sel = gdb.make_selector(gdb.SEL_RANDOM, narrow_groups=2, seed=100 + runs, trace=trace,
                                        trace_count=cnt)
                solve_one(ctx, g, dg, algo, sel)
                runs += 1
        bounds.append(min(int(cnt.to_numpy()[0]), cap))
    raw = trace.to_numpy().reshape(cap, gdb.TRACE_STRIDE)
    rec = balanced([steady_records(raw[bounds[i]:bounds[i + 1]]) for i in range(3)])
    drec = gdb.DeviceBuffer.from_numpy(rec)
    p, w, loss = gdb.mlp_train(ctx, gdb.mlp_init(3), drec, len(rec), steps=2000, batch=256, lr=2e-3, seed=3)
    print("records", len(rec), "loss", loss)
    gains = {}
    for name, g in (("rmat16", gdb.gen_rmat(16, 16, seed=61, wlo=1, whi=255)),
                    ("rmat17", gdb.gen_rmat(17, 16, seed=62, wlo=1, whi=255))):
        dg = gdb.DeviceCSR(g)
        for algo in (gdb.TASK_BFS, gdb.TASK_SSSP):
            ref, t_push, _ = solve_one(ctx, g, dg, algo, None)
            got, t_mlp, arms = solve_one(ctx, g, dg, algo, gdb.make_selector(gdb.SEL_MLP, narrow_groups=2,
                                                                            weights=w))
            assert (got == ref).all()
            gains[(name, algo)] = (t_push, t_mlp, arms)
    print(gains)
    bfs_push = sum(v[0] for k, v in gains.items() if k[1] == gdb.TASK_BFS)
    bfs_mlp = sum(v[1] for k, v in gains.items() if k[1] == gdb.TASK_BFS)
    tot_push = sum(v[0] for v in gains.values())
    tot_mlp = sum(v[1] for v in gains.values())
    print(f"BFS gain {bfs_push / bfs_mlp:.2f}x, all {tot_push / tot_mlp:.2f}x")
    assert bfs_mlp < 0.7 * bfs_push
    assert tot_mlp < tot_push
