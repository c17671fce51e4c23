"""Headline benchmark: GTEPS of the worklist BFS + SSSP solve on a 2048x2048
4-neighbour grid (5% of edges removed, seed 2, weights U[1,255], source 0),
on one MI355X.

A step is one BFS solve plus one SSSP solve with the graph resident in HBM,
both in one persistent launch of the batched engine (two teams side by side;
the engine and shape the tuning table names: owner-partitioned vertex ranges
in LDS when the autotuner found them faster).  value = 2 * m_r / (kernel
time), with m_r = sum of out-degrees of reached vertices.  The plain
persistent engine (the reference kernels' structure, default launch) and the
tuned single-task kernels run on the same input; all outputs are compared bit
for bit and checked with dverifysolution.  Also reported: us per barrier vs
us per relaunch, the roofline of the dominant kernel with its measured HBM
traffic, and a CPU baseline (the oracle's threaded restatement, and its
single-thread time beside it).

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--config c2]
(--config c1 / c3 / c4 measure the other configurations the same way.)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
GRID_SIDE = 2048


def tuned_c2(engine: str = ""):
    """The batched engine's shape for the grid from the library's tuning table
    (gpu_discovery_barrier_amd/tune_table.txt, written by the autotuner):
    ((wgs, groups, band, local hops, selector mode), team groups, LDS queue
    tile, owned).  The autotuner writes a "batch_owned" row only when the
    owner-partitioned engine beat the shared one on this grid; that row then
    names the headline engine.  engine = "owned" / "shared" forces one."""
    from gpu_discovery_barrier_amd import autotune

    n = GRID_SIDE * GRID_SIDE
    r = autotune.lookup("batch_owned", n) if engine != "shared" else None
    if r is not None or engine == "owned":
        if r is None:
            return (1024, 256, 6144, 0, 0), 128, 2048, True
        return ((r["wgs"], r["groups"], r["delta"], 0, 0), r["team_groups"] or max(1, r["groups"] // 2),
                r["lds_queue"] or 2048, True)
    r = autotune.lookup("batch", n)
    if r is None:
        return (128, 256, 4096, 256, 0), 128, 2048, False
    return ((r["wgs"], r["groups"], r["delta"], r["hops"], 0), r["team_groups"] or max(1, r["groups"] // 2),
            r["lds_queue"] or 2048, False)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def reached_edges(g, dist, inf) -> int:
    deg = np.diff(g.row).astype(np.int64)
    return int(deg[dist != inf].sum())


def traffic_from_profiles(shape: tuple):
    """HBM bytes per launch of the batched kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the same
    counters' reading for a copy kernel of known size), when it was taken at
    this launch shape; else None."""
    path = ROOT / "profiles" / "batch_grid2048_traffic.json"
    try:
        d = json.loads(path.read_text())
        if tuple(d["shape"]) != tuple(shape):
            return None
        return d["hbm_bytes_per_launch"]
    except Exception:
        return None


def cpu_baseline(g, want_b=None, want_s=None) -> dict:
    """The worklist BFS + SSSP on the host's cores: persistent threads (one per
    core, at most 16, the GPU box's CPU share) running the same iterations
    with a CPU XF barrier; one warm-up, then the median of 5 solves, timed
    around the solve only.  The single-thread sequential restatement is timed
    beside it."""
    import oracle

    threads = min(16, os.cpu_count() or 1)
    tb, ts = [], []
    for rep in range(6):
        t0 = time.perf_counter()
        dist_b, _ = oracle.worklist_mt(g.row, g.col, None, 0, threads=threads)
        t1 = time.perf_counter()
        dist_s, _ = oracle.worklist_mt(g.row, g.col, g.w, 0, threads=threads)
        t2 = time.perf_counter()
        if rep > 0:
            tb.append(t1 - t0)
            ts.append(t2 - t1)
    exact = (want_b is None or bool((dist_b == want_b).all())) and (want_s is None or bool((dist_s == want_s).all()))
    mr = reached_edges(g, dist_b, oracle.LONESTAR_INF) + reached_edges(g, dist_s, oracle.LONESTAR_INF)
    t_mt = float(np.median(tb)) + float(np.median(ts))
    t0 = time.perf_counter()
    oracle.bfs_worklist_ref(g.row, g.col, 0)
    t1 = time.perf_counter()
    oracle.sssp_worklist_ref(g.row, g.col, g.w, 0)
    t2 = time.perf_counter()
    return {
        "value": round(mr / t_mt / 1e9, 6),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"one worklist BFS + one worklist SSSP of the full {GRID_SIDE}x{GRID_SIDE} grid on {threads} "
                  f"persistent threads with a CPU XF barrier (oracle/oracle.c orc_worklist_mt), median of 5 after "
                  f"one warm-up: bfs {np.median(tb):.3f}s, sssp {np.median(ts):.3f}s; outputs equal the GPU's: {exact}",
        "single_thread": {"value": round(mr / (t2 - t0) / 1e9, 6), "bfs_s": round(t1 - t0, 3),
                          "sssp_s": round(t2 - t1, 3), "kind": "sequential restatement"},
    }


def barrier_table(ctx, wgs: int, group_counts, K: int = 2000) -> dict:
    """us per barrier for XF / counter / tree at each group count, and us per
    back-to-back relaunch of an empty kernel of that many groups:
    (t(K barriers) - t(0 barriers)) / K inside one persistent kernel."""
    import gpu_discovery_barrier_amd as gdb

    cap = gdb.max_resident_groups(wgs, 0)
    out = {"wgs": wgs, "K": K}
    for kind, name in ((gdb.BARRIER_XF, "xf"), (gdb.BARRIER_COUNTER, "counter"), (gdb.BARRIER_TREE, "tree")):
        ctx.configure(gdb.PROTOCOL_LOCKFREE, kind)
        out[name] = {}
        for g in group_counts:
            g = min(g, cap)
            t0, _ = ctx.barrier_bench(wgs, g, 0)
            tk, p = ctx.barrier_bench(wgs, g, K)
            out[name][str(g)] = {"us_per_barrier": round((tk - t0) * 1e3 / K, 3), "P": p}
    ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)
    out["relaunch"] = {str(min(g, cap)): round(ctx.relaunch_bench(wgs, min(g, cap), K, 0) * 1e3 / K, 3)
                       for g in group_counts}
    return out


def line(metric, value, unit, args, ms_per_step, config, dtype, extra) -> dict:
    out = {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def timed(fn, warmup, steps):
    """Runs fn() -> kernel ms; returns (median kernel ms, median wall ms, last result)."""
    for _ in range(warmup):
        fn()
    ks, ws, res = [], [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        res = fn()
        ws.append((time.perf_counter() - t0) * 1e3)
        ks.append(res[0])
    return float(np.median(ks)), float(np.median(ws)), res


def bench_c1(args) -> int:
    """Pannotia max-rule colouring, R-MAT scale 14 (edge factor 16, seed 1),
    priorities a seeded permutation.  Persistent engine (autotuned launch
    shape, one barrier per round) vs the plain persistent engine (the
    reference kernel's structure, default launch).  GTEPS = m / time."""
    import gpu_discovery_barrier_amd as gdb
    import oracle

    g = gdb.gen_rmat(14, 16, seed=1)
    dg = gdb.DeviceCSR(g, weights=False)
    nv_h = (gdb.permutation(g.n, 1).astype(np.float64) / (g.n + 1)).astype(np.float32)
    nv = gdb.DeviceBuffer.from_numpy(nv_h)
    ctx = gdb.Context()
    out = gdb.DeviceBuffer.empty(g.n, np.int32)

    def run(variant, wgs, groups):
        c, st = gdb.color(ctx, dg, nv, gdb.Opts.make(wgs=wgs, groups=groups, variant=variant), out=out)
        return st.kernel_ms, c, st

    trials = []
    for w in (64, 128, 256, 512, 1024):
        for gr in (8, 16, 32, 64, 128, 256, 512):
            if gr <= gdb.max_resident_groups(w, 0):
                trials.append((min(run(gdb.VARIANT_TUNED, w, gr)[0] for _ in range(3)), w, gr))
    trials.sort()
    _, wgs, groups = trials[0]
    t_ms, wall, (_, c_t, st_t) = timed(lambda: run(gdb.VARIANT_TUNED, wgs, groups), args.warmup, args.steps)
    p_ms, _, (_, c_p, st_p) = timed(lambda: run(gdb.VARIANT_PLAIN, 256, 0), 1, 3)
    want, rounds = oracle.color_ref(g.row, g.col, nv_h)
    exact = bool((c_t == want).all() and (c_p == want).all())
    t0 = time.perf_counter()
    oracle.color_ref(g.row, g.col, nv_h)
    t_cpu = time.perf_counter() - t0
    gteps = g.m / (t_ms / 1e3) / 1e9
    o = line("GTEPS on 1 MI355X (persistent autotuned vs plain, bit-exact)", gteps, "GTEPS", args, wall,
             {"workload": f"Pannotia max-rule colour, R-MAT scale 14 ef 16 seed 1 (n={g.n}, m={g.m}), "
                          "priorities = seeded permutation",
              "wgs": wgs, "groups": groups, "participating": st_t.participating, "rounds": st_t.rounds},
             "f32", {
                 "kernel_ms": round(t_ms, 4), "plain": {"gteps": round(g.m / (p_ms / 1e3) / 1e9, 4),
                                                        "ms": round(p_ms, 4), "participating": st_p.participating,
                                                        "barriers": st_p.barriers},
                 "speedup_vs_plain": round(p_ms / t_ms, 3), "bit_exact_vs_plain": exact,
                 "oracle_rounds": rounds,
                 "cpu_baseline": {"value": round(g.m / t_cpu / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
                                  "sample": f"one full colouring by the oracle restatement ({t_cpu:.3f}s)"}})
    ctx.close()
    print(json.dumps(o), flush=True)
    return 0 if exact else 3


def bench_c3(args) -> int:
    """Pannotia MIS and BC on R-MAT scale 22 (edge factor 16, seed 3).  MIS
    priorities: a seeded permutation; BC: 64 seeded sources, f32 sigma /
    delta.  Tuned persistent kernels (autotuned launch) vs the plain
    persistent kernels.  value = GTEPS of MIS + BC together:
    (m + 64 m_r) / (t_mis + t_bc)."""
    import gpu_discovery_barrier_amd as gdb
    import oracle

    log("generating R-MAT 22 ...")
    g = gdb.gen_rmat(22, 16, seed=3)
    dg = gdb.DeviceCSR(g, weights=False)
    nv_h = (gdb.permutation(g.n, 3).astype(np.float64) / (g.n + 1)).astype(np.float32)
    nv = gdb.DeviceBuffer.from_numpy(nv_h)
    deg = np.diff(g.row)
    rng = np.random.default_rng(3)
    sources = np.sort(rng.choice(np.flatnonzero(deg > 0), 64, replace=False)).astype(np.int32)
    ctx = gdb.Context()
    mis_out = gdb.DeviceBuffer.empty(g.n, np.int32)
    bc_out = gdb.DeviceBuffer.empty(g.n, np.float32)

    def mis(variant, wgs, groups):
        s, st = gdb.mis(ctx, dg, nv, gdb.Opts.make(wgs=wgs, groups=groups, variant=variant), out=mis_out)
        return st.kernel_ms, s, st

    def bc(variant, wgs, groups):
        b, st = gdb.bc(ctx, dg, dg, sources, gdb.Opts.make(wgs=wgs, groups=groups, variant=variant), out=bc_out)
        return st.kernel_ms, b, st

    if args.engine_only:
        # Tuned kernels only, at the tuning table's shapes (profiler runs).
        from gpu_discovery_barrier_amd import autotune

        rm, rb = autotune.lookup("mis", g.n), autotune.lookup("bc", g.n)
        sm = (rm["wgs"], rm["groups"]) if rm else (256, 2048)
        sb = (rb["wgs"], rb["groups"]) if rb else (256, 2048)
        m_ms = mis(gdb.VARIANT_TUNED, *sm)[0]
        b_ms = bc(gdb.VARIANT_TUNED, *sb)[0]
        print(json.dumps({"mis_ms": round(m_ms, 3), "bc_ms": round(b_ms, 3), "mis_shape": sm, "bc_shape": sb}),
              flush=True)
        ctx.close()
        return 0
    shapes = [(w, gr) for w in (256, 512, 1024) for gr in (256, 512, 1024, 2048)
              if gr <= gdb.max_resident_groups(w, 0)]
    mt = sorted((mis(gdb.VARIANT_TUNED, w, gr)[0], w, gr) for w, gr in shapes)[0]
    # BC runs take ~0.6 s each: a few shapes around the usual pick.
    bshapes = [(w, gr) for w, gr in ((256, 1024), (256, 2048), (512, 512), (512, 1024))
               if gr <= gdb.max_resident_groups(w, 0)]
    bt = sorted((bc(gdb.VARIANT_TUNED, w, gr)[0], w, gr) for w, gr in bshapes)[0]
    log(f"autotune: mis wgs={mt[1]} groups={mt[2]}, bc wgs={bt[1]} groups={bt[2]}")
    m_ms, wall_m, (_, s_t, st_m) = timed(lambda: mis(gdb.VARIANT_TUNED, mt[1], mt[2]), args.warmup, args.steps)
    b_ms, wall_b, (_, b_t, st_b) = timed(lambda: bc(gdb.VARIANT_TUNED, bt[1], bt[2]), 1, max(1, args.steps // 2))
    pm_ms, _, (_, s_p, _) = timed(lambda: mis(gdb.VARIANT_PLAIN, 256, 0), 0, 1)
    pb_ms, _, (_, b_p, _) = timed(lambda: bc(gdb.VARIANT_PLAIN, 256, 0), 0, 1)
    # m_r per source: edges of the vertices each BFS reaches.
    lv = gdb.DeviceBuffer.empty(g.n, np.uint32)
    mr = 0
    for s0 in sources:
        d, _ = gdb.bfs_worklist(ctx, dg, int(s0), gdb.Opts.make(variant=gdb.VARIANT_TUNED, wgs=512, groups=256), out=lv)
        mr += int(deg[d != gdb.LONESTAR_INF].sum())
    t0 = time.perf_counter()
    want = oracle.mis_tb(g.row, g.col, nv_h)
    t_cpu = time.perf_counter() - t0
    mis_exact = bool((s_t == want).all() and (s_p == want).all())
    # BC against FP64 Brandes over the same 64 sources, entry by entry:
    # |got - want| <= 1e-4 |want| (floor 1e-6 for zero entries).
    log("FP64 Brandes oracle (threads) ...")
    bc_want = oracle.bc_brandes_mt(g.row, g.col, sources)
    nz = bc_want != 0

    def bc_check(b):
        err = np.abs(b.astype(np.float64) - bc_want)
        ok = bool((err <= np.maximum(1e-4 * np.abs(bc_want), 1e-6)).all())
        return ok, float((err[nz] / np.abs(bc_want[nz])).max())

    bc_ok_t, bc_rel = bc_check(b_t)
    bc_ok_p, bc_rel_p = bc_check(b_p)
    # The same BC on the graph relabelled by decreasing degree (the hubs every
    # traversal touches share the first lines of each per-vertex array),
    # results mapped back to the original ids; the relabelling is host-side
    # preprocessing, timed on its own.
    t0 = time.perf_counter()
    nid = g.order_by_degree()
    gr = g.relabel(nid)
    t_relabel = time.perf_counter() - t0
    dgr = gdb.DeviceCSR(gr, weights=False)
    src_r = nid[sources].astype(np.int32)

    def bc_relabelled():
        b, st = gdb.bc(ctx, dgr, dgr, src_r, gdb.Opts.make(wgs=bt[1], groups=bt[2], variant=gdb.VARIANT_TUNED),
                       out=bc_out)
        return st.kernel_ms, b, st

    br_ms, _, (_, b_rn, _) = timed(bc_relabelled, 1, max(1, args.steps // 2))
    bc_ok_r, bc_rel_r = bc_check(b_rn[nid])
    del dgr
    total_ms = m_ms + b_ms
    gteps = (g.m + mr) / (total_ms / 1e3) / 1e9
    o = line("GTEPS on 1 MI355X (tuned persistent vs plain; MIS bit-exact, BC f32 within 1e-4 of FP64 per entry)", gteps, "GTEPS",
             args, wall_m + wall_b,
             {"workload": f"Pannotia MIS + BC (64 sources), R-MAT scale 22 ef 16 seed 3 (n={g.n}, m={g.m})",
              "mis": {"wgs": mt[1], "groups": mt[2], "rounds": st_m.rounds},
              "bc": {"wgs": bt[1], "groups": bt[2], "sources": 64}},
             "f32", {
                 "mis": {"ms": round(m_ms, 3), "gteps": round(g.m / m_ms / 1e6, 4), "plain_ms": round(pm_ms, 3),
                         "speedup_vs_plain": round(pm_ms / m_ms, 3), "bit_exact": mis_exact},
                 "bc": {"ms": round(b_ms, 3), "gteps": round(mr / b_ms / 1e6, 4), "plain_ms": round(pb_ms, 3),
                        "speedup_vs_plain": round(pb_ms / b_ms, 3), "max_rel_err_vs_fp64_oracle": bc_rel, "within_1e-4": bc_ok_t,
                        "plain_max_rel_err_vs_fp64_oracle": bc_rel_p, "plain_within_1e-4": bc_ok_p,
                        "m_r_sum": mr,
                        "degree_relabelled": {"ms": round(br_ms, 3), "gteps": round(mr / br_ms / 1e6, 4),
                                              "max_rel_err_vs_fp64_oracle": bc_rel_r, "within_1e-4": bc_ok_r,
                                              "relabel_host_s": round(t_relabel, 3)}},
                 "speedup_vs_plain": round((pm_ms + pb_ms) / total_ms, 3),
                 "cpu_baseline": {"value": round(g.m / t_cpu / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
                                  "sample": f"greedy sequential MIS of the same graph (oracle, {t_cpu:.2f}s)"}})
    ctx.close()
    print(json.dumps(o), flush=True)
    return 0 if mis_exact and bc_ok_t and bc_ok_p and bc_ok_r else 3


def bench_c4(args) -> int:
    """Batched: 4096 graphs (2048 G(n,p), n log-uniform in [256, 65536],
    average degree 8; 2048 grids, side uniform in [16, 256]; weights in
    [1, 100]; seeds 4 + graph index), each solved with BFS, SSSP and colour:
    12288 tasks in one persistent launch vs the plain persistent engine one
    task per launch (timed on every 16th graph; its GTEPS is the sample's).
    GTEPS = sum of m_r (BFS, SSSP) and m (colour) / time."""
    import gpu_discovery_barrier_amd as gdb
    import oracle

    rng = np.random.default_rng(4)
    log("generating 4096 graphs ...")
    gs = []
    for i in range(2048):
        n = int(np.floor(np.exp(rng.uniform(np.log(256), np.log(65536)))))
        gs.append(gdb.gen_gnp(n, 8.0, seed=4 + i, wlo=1, whi=100))
    for i in range(2048):
        sd = int(rng.integers(16, 257))
        gs.append(gdb.gen_grid(sd, sd, 0.0, seed=4 + 2048 + i, wlo=1, whi=100))
    log("uploading ...")
    dgs = [gdb.DeviceCSR(g) for g in gs]
    nvs = [gdb.DeviceBuffer.from_numpy(gdb.pannotia_values(g.n)) for g in gs]
    outs = [[gdb.DeviceBuffer.empty(g.n, np.uint32), gdb.DeviceBuffer.empty(g.n, np.uint32),
             gdb.DeviceBuffer.empty(g.n, np.int32)] for g in gs]
    ctx = gdb.Context()
    N = sum(g.n for g in gs)
    M = sum(g.m for g in gs)

    def tasks(delta):
        ts = []
        for dg, nv, o in zip(dgs, nvs, outs):
            ts.append(gdb.make_task(gdb.TASK_BFS, dg, o[0], 0))
            ts.append(gdb.make_task(gdb.TASK_SSSP, dg, o[1], 0, delta=delta))
            ts.append(gdb.make_task(gdb.TASK_COLOR, dg, o[2], 0, node_value=nv))
        return ts

    def run(wgs, groups, team, delta):
        ts = tasks(delta)
        _, st = gdb.batch_solve(ctx, ts, team, gdb.Opts.make(wgs=wgs, groups=groups))
        return st.kernel_ms, st

    trials = []
    for w in (128, 256, 512):
        for gr in (512, 1024, 2048):
            if gr > gdb.max_resident_groups(w, 0):
                continue
            for dl in (0, 200):
                trials.append((run(w, gr, 1, dl)[0], w, gr, dl))
    trials.sort()
    _, wgs, groups, delta = trials[0]
    log(f"autotune: wgs={wgs} groups={groups} delta={delta} ({trials[0][0]:.1f} ms)")
    k_ms, wall, (_, st) = timed(lambda: run(wgs, groups, 1, delta), args.warmup, args.steps)
    # The same batch captured once as a HIP graph and replayed.
    gtasks = tasks(delta)
    graph = gdb.BatchGraph(ctx, gtasks, 1, gdb.Opts.make(wgs=wgs, groups=groups))
    g_ms, g_wall, _ = timed(lambda: (graph.launch()[1].kernel_ms, None), args.warmup, args.steps)
    graph.close()
    # Reached edges (BFS and SSSP reach the same vertices) and parity on a sample.
    deg_r = []
    exact = True
    plain_ms, plain_edges = 0.0, 0
    for i, (g, dg, nv, o) in enumerate(zip(gs, dgs, nvs, outs)):
        lv = o[0].to_numpy()
        mr = int(np.diff(g.row)[lv != gdb.LONESTAR_INF].sum())
        deg_r.append(mr)
        if i % 16 == 0:
            want_b, _ = oracle.bfs_worklist_ref(g.row, g.col, 0)
            want_s = oracle.dijkstra_tb(g.row, g.col, g.w, 0)
            want_c, _ = oracle.color_ref(g.row, g.col, gdb.pannotia_values(g.n))
            exact &= bool((lv == want_b).all() and (o[1].to_numpy().astype(np.int64) == want_s).all()
                          and (o[2].to_numpy() == want_c).all())
            po = gdb.Opts.make(variant=gdb.VARIANT_PLAIN)
            _, s1 = gdb.bfs_worklist(ctx, dg, 0, po)
            _, s2 = gdb.sssp_worklist(ctx, dg, 0, po)
            _, s3 = gdb.color(ctx, dg, nv, po)
            plain_ms += s1.kernel_ms + s2.kernel_ms + s3.kernel_ms
            plain_edges += 2 * mr + g.m
    work = 2 * sum(deg_r) + M
    gteps = work / (k_ms / 1e3) / 1e9
    plain_gteps = plain_edges / (plain_ms / 1e3) / 1e9
    o = line("GTEPS on 1 MI355X (batched vs plain, bit-exact)", gteps, "GTEPS", args, wall,
             {"workload": f"4096 graphs (2048 G(n,p) + 2048 grids, n={N}, m={M}) x BFS, SSSP, colour",
              "wgs": wgs, "groups": groups, "team_groups": 1, "delta": delta, "participating": st.participating},
             "u32", {"kernel_ms": round(k_ms, 3), "work_edges": work,
                     "graph_replay": {"ms": round(g_ms, 3), "wall_ms": round(g_wall, 3), "eager_wall_ms": round(wall, 3)},
                     "plain": {"gteps": round(plain_gteps, 4), "sample": "every 16th graph, one task per launch",
                               "ms_sample": round(plain_ms, 3)},
                     "speedup_vs_plain": round(gteps / plain_gteps, 3), "bit_exact_sample": exact})
    ctx.close()
    print(json.dumps(o), flush=True)
    return 0 if exact else 3


def bench_probe(args) -> int:
    """Memory-latency probes and protocol / barrier costs on this device:
    dependent-access chains (plain load, agent-scope load, returning atomic)
    over buffers served by L2, the Infinity Cache and HBM, in coarse- and
    fine-grained memory; dependent atomic-min steps (32- vs 64-bit) from one
    CU and from 256 CUs; the discovery protocol's cost per lock kind for 1000
    launched groups; every barrier kind at 8..1024 groups for wgs 128 / 256."""
    import gpu_discovery_barrier_amd as gdb

    ctx = gdb.Context()
    steps = 20000
    chain = {}
    for nbytes, label in ((1 << 20, "1MiB"), (16 << 20, "16MiB"), (128 << 20, "128MiB"), (1 << 30, "1GiB")):
        for kind, kn in ((0, "load"), (2, "agent_load"), (1, "atomic")):
            chain[f"{kn}_{label}"] = round(ctx.probe_chain(kind, nbytes, False, steps), 1)
        log(f"chain {label}: {[chain[k] for k in chain if k.endswith(label)]}")
    for kind, kn in ((0, "load"), (2, "agent_load"), (1, "atomic")):
        chain[f"{kn}_1MiB_fine"] = round(ctx.probe_chain(kind, 1 << 20, True, steps), 1)
    atom = {}
    for bits in (32, 64):
        for groups, wgs in ((1, 1024), (256, 1024), (256, 128)):
            atom[f"{bits}bit_{groups}x{wgs}"] = round(ctx.probe_atomics(bits, groups, wgs, 2000), 1)
    ping = {}
    for mode, mn in ((0, "relaxed"), (1, "release_acquire"), (2, "fetch_add_acq_rel")):
        for fine in (False, True):
            for same in (True, False):
                ns, xc = ctx.probe_pingpong(mode, fine, same, 5000)
                ping[f"{mn}_{'fine' if fine else 'coarse'}_{'same_xcd' if same else 'two_xcds'}"] = {
                    "ns": round(ns, 1), "xcds": xc}
    log(f"pingpong: {ping}")
    prot = {}
    for lock, name in ((gdb.PROTOCOL_LOCKFREE, "lock-free"), (gdb.PROTOCOL_TICKET, "ticket"),
                       (gdb.PROTOCOL_SPIN, "spin")):
        ctx.configure(lock, gdb.BARRIER_AUTO)
        for wgs in (64, 256):
            ns = ctx.time_protocol(1000, wgs)
            prot[f"{name}_wgs{wgs}"] = {"ms": round(ns / 1e6, 4), "P": ctx.participating()}
    ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)
    bars = {str(w): barrier_table(ctx, w, [8, 32, 128, 256, 1024]) for w in (128, 256)}
    out = {"metric": "memory latency and synchronisation costs on 1 MI355X", "unit": "ns / us",
           "chain_ns_per_access": chain, "atomic_ns_per_dependent_step": atom, "pingpong_ns_per_handoff": ping,
           "protocol_1000_groups": prot, "barrier_us": bars, "steps": steps, "data": "synthetic"}
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


def bench_sched(args) -> int:
    """Persistent barrier vs kernel relaunch as the schedule of a worklist
    solve: BFS and SSSP on the 2048^2 grid (thousands of short iterations)
    and on R-MAT 20 (a dozen wide ones), each run (a) persistent tuned, one
    inter-group barrier per iteration, and (b) relaunched per iteration with
    the host reading the stop flag after every 1, 4 or 16 launches (the
    reference's color.cpp:294-334 loop and its batched variant).  Reported:
    kernel ms and us per iteration; outputs compared bit for bit."""
    import gpu_discovery_barrier_amd as gdb

    ctx = gdb.Context()
    rows = []
    for name, g in (("grid2048", gdb.gen_grid(GRID_SIDE, GRID_SIDE, 0.05, seed=2, wlo=1, whi=255)),
                    ("rmat20", gdb.gen_rmat(20, 16, seed=50, wlo=1, whi=255))):
        dg = gdb.DeviceCSR(g)
        for algo, fn in (("bfs", gdb.bfs_worklist), ("sssp", gdb.sssp_worklist)):
            ref, st = fn(ctx, dg, 0, gdb.Opts.make(variant=gdb.VARIANT_TUNED, wgs=512, groups=128))
            ms = min(fn(ctx, dg, 0, gdb.Opts.make(variant=gdb.VARIANT_TUNED, wgs=512, groups=128))[1].kernel_ms
                     for _ in range(2))
            r = {"graph": name, "algo": algo, "iterations": st.rounds,
                 "persistent": {"ms": round(ms, 3), "us_per_iteration": round(ms * 1e3 / max(1, st.rounds), 3)}}
            for check in (1, 4, 16):
                o = gdb.Opts.make(schedule=gdb.SCHED_RELAUNCH, wgs=256, relaunch_check=check)
                got, s2 = fn(ctx, dg, 0, o)
                r[f"relaunch_check{check}"] = {"ms": round(s2.kernel_ms, 3), "launches": s2.launches,
                                               "us_per_iteration": round(s2.kernel_ms * 1e3 / max(1, s2.launches), 3),
                                               "exact": bool((got == ref).all())}
            log(json.dumps(r))
            rows.append(r)
    wins = [r for r in rows if min(r[k]["ms"] for k in r if k.startswith("relaunch")) < r["persistent"]["ms"]]
    out = {"metric": "persistent barrier vs relaunch schedule, worklist BFS / SSSP on 1 MI355X", "unit": "ms",
           "rows": rows, "relaunch_wins": len(wins), "data": "synthetic"}
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


def copy_calibration(args) -> int:
    """Three device-to-device copies of 256 MiB (the runtime's copy kernel):
    the reference reading for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import gpu_discovery_barrier_amd as gdb

    nbytes = 256 << 20
    a = gdb.DeviceBuffer(nbytes, np.uint8)
    b = gdb.DeviceBuffer(nbytes, np.uint8)
    for _ in range(3):
        gdb.check(gdb.lib().gdb_memcpy_dtod(b.ptr, a.ptr, nbytes), "gdb_memcpy_dtod")
    gdb.check(gdb.lib().gdb_device_synchronize(), "gdb_device_synchronize")
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--tune", action="store_true",
                    help="c2: rerun the autotuner (launch shape, band, hops, selector) instead of the tuned shape")
    ap.add_argument("--shape", default="",
                    help="c2: wgs,groups,delta,hops,selector (skips the autotuner; selector 0 none, 4 MLP)")
    ap.add_argument("--engine", default="", choices=["", "owned", "shared"],
                    help="c2: owner-partitioned ranges or the shared engine (default: the tuning table's pick)")
    ap.add_argument("--engine-only", action="store_true",
                    help="c2 / c3: time the engine's kernels only (profiler runs)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "copy", "probe", "sched"],
                    help="c2 (default): the headline grid BFS + SSSP; c1 colour R-MAT 14; "
                         "c3 MIS + BC R-MAT 22; c4 4096-graph batch; copy: counter calibration; "
                         "probe: memory latency, protocol and barrier costs; sched: persistent vs relaunch")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build runs a single MI355X)", file=sys.stderr)
        return 2
    if args.config != "c2":
        return {"c1": bench_c1, "c3": bench_c3, "c4": bench_c4, "copy": copy_calibration,
                "probe": bench_probe, "sched": bench_sched}[args.config](args)

    import gpu_discovery_barrier_amd as gdb
    from gpu_discovery_barrier_amd import autotune

    if gdb.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    log("generating grid ...")
    g = gdb.gen_grid(GRID_SIDE, GRID_SIDE, 0.05, seed=2, wlo=1, whi=255)
    dg = gdb.DeviceCSR(g)
    dist = gdb.DeviceBuffer.empty(g.n, np.uint32)
    ctx = gdb.Context()

    def solve(variant, wgs=256, groups=0):
        o = gdb.Opts.make(wgs=wgs, groups=groups, variant=variant)
        db, sb = gdb.bfs_worklist(ctx, dg, 0, o, out=dist)
        ds, ss = gdb.sssp_worklist(ctx, dg, 0, o, out=dist)
        return db, sb, ds, ss

    out_b = gdb.DeviceBuffer.empty(g.n, np.uint32)
    out_s = gdb.DeviceBuffer.empty(g.n, np.uint32)
    last_tasks = []

    # The learned schedule selector: the MLP trained on the engine's own
    # traces (gpu_discovery_barrier_amd/selector_weights.npy).
    wpath = ROOT / "gpu_discovery_barrier_amd" / "selector_weights.npy"
    weights = gdb.DeviceBuffer.from_numpy(np.load(wpath).astype(np.uint16)) if wpath.exists() else None
    narrow = 2

    def selector(mode):
        if mode == gdb.SEL_NONE:
            return None
        return gdb.make_selector(mode, narrow_groups=narrow, weights=weights)

    tuned_shape, team, qcap, owned = tuned_c2(args.engine)
    if args.tune or args.shape:
        owned = args.engine == "owned"

    def solve_batched(wgs, groups, delta, hops=0, mode=gdb.SEL_NONE, team_groups=None, lds_queue=None):
        # One launch, two teams: BFS and SSSP side by side.  The grid is
        # symmetric, so it is its own in-edge CSR for pull arms.  Owned:
        # every group of a team holds a contiguous vertex range in LDS.
        tasks = [gdb.make_task(gdb.TASK_BFS, dg, out_b, 0, hops=hops, transpose=dg, owned=owned),
                 gdb.make_task(gdb.TASK_SSSP, dg, out_s, 0, delta=delta, hops=hops, transpose=dg, owned=owned)]
        rounds, st = gdb.batch_solve(ctx, tasks, team_groups or team,
                                     gdb.Opts.make(wgs=wgs, groups=groups, lds_queue=lds_queue or qcap),
                                     selector=selector(mode))
        last_tasks[:] = tasks
        return rounds, st

    # The shape: the tuning table's row for this workload (written by
    # python -m gpu_discovery_barrier_amd.autotune, which sweeps wgs, groups,
    # team, band, local hops and the LDS queue tile), or --shape, or --tune
    # (reruns that sweep here, plus the learned selector, which needs hops 0).
    if args.shape:
        wgs, groups, delta, hops, mode = (int(x) for x in args.shape.split(","))
        team = max(1, groups // 2)
    elif not args.tune:
        wgs, groups, delta, hops, mode = tuned_shape
    else:
        from gpu_discovery_barrier_amd import autotune

        rows, trials = autotune.tune_batch_grid(ctx)
        r = rows[-1]  # the owned row when it won
        owned = r["solver"] == "batch_owned"
        wgs, groups, team, delta, hops, qcap = r["wgs"], r["groups"], r["team_groups"], r["delta"], r["hops"], \
            r["lds_queue"]
        mode = gdb.SEL_NONE
        if weights is not None and not owned:  # the selector's arms belong to the shared engine
            ms_mlp = min(solve_batched(wgs, groups, delta, 0, gdb.SEL_MLP)[1].kernel_ms for _ in range(2))
            if ms_mlp < min(t[0] for t in trials):
                hops, mode = 0, gdb.SEL_MLP
        log(f"autotune: wgs={wgs} groups={groups} team={team} delta={delta} hops={hops} lds_queue={qcap} "
            f"selector={mode}; {len(trials)} shapes tried")

    for _ in range(args.warmup):
        solve_batched(wgs, groups, delta, hops, mode)
    step_ms, walls = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        rounds, bst = solve_batched(wgs, groups, delta, hops, mode)
        walls.append(time.perf_counter() - t0)
        step_ms.append(bst.kernel_ms)
    db, ds = out_b.to_numpy(), out_s.to_numpy()

    inf = gdb.LONESTAR_INF
    mr = reached_edges(g, db, inf)
    t_step_ms = float(np.median(step_ms))
    gteps = 2 * mr / (t_step_ms / 1e3) / 1e9
    if args.engine_only:
        # Only the batched launches (for profiler runs): no comparisons.
        print(json.dumps({"kernel_ms": round(t_step_ms, 3), "gteps": round(gteps, 4),
                          "shape": [wgs, groups, delta, hops, mode, int(owned)],
                          "work_edges": bst.work_edges, "work_items": bst.work_items}), flush=True)
        ctx.close()
        return 0

    # Single-task tuned kernels (one solve per launch) at their own best shape.
    tb, tsb, tds, tss = solve(gdb.VARIANT_TUNED, 512, 128)
    tuned_ms = tsb.kernel_ms + tss.kernel_ms

    # Plain engine (reference kernel structure, default launch shape).
    pdb, psb, pds, pss = solve(gdb.VARIANT_PLAIN, 256, 0)
    plain_ms = psb.kernel_ms + pss.kernel_ms
    plain_gteps = 2 * mr / (plain_ms / 1e3) / 1e9
    bit_exact = bool((pdb == db).all() and (pds == ds).all() and (tb == db).all() and (tds == ds).all())
    nerr = gdb.verify_distances(ctx, dg, gdb.DeviceBuffer.from_numpy(ds), weighted=True)
    nerr += gdb.verify_distances(ctx, dg, gdb.DeviceBuffer.from_numpy(db), weighted=False)

    # Barrier vs relaunch: every barrier kind at the team's group count (one
    # team = groups / 2) and at 8, 32, 256 and 1024 groups, all of 128
    # threads (the shared engine's team shape; the owned engine's rounds use
    # workgroup barriers, and its team barrier only at set-up and band ends);
    # relaunch (i) back to back, (ii) with the reference host loop's blocking
    # readback.
    P, bwgs = team, 128
    bars = barrier_table(ctx, bwgs, sorted({8, 32, P, 256, 1024}))
    us_bar = bars["counter" if P < 64 else "tree"][str(P)]["us_per_barrier"]
    us_relaunch = bars["relaunch"][str(P)]
    us_relaunch_sync = ctx.relaunch_bench(bwgs, P, 500, 1) * 1e3 / 500

    # Roofline of the dominant kernel (the batched kernel).  Algorithmic bytes
    # as the kernels count their work: BFS 8 B per processed edge (dst, dist)
    # + 24 B per processed item; SSSP 16 B per processed edge (dst, weight,
    # dist[src], dist[dst]) + 24 B per processed item.  Minimal bytes charge
    # every reached edge and vertex once (m_r, n_r), as a level-synchronous
    # solve with no re-relaxation would.
    tb, ts = last_tasks
    alg_bytes = 8 * tb.work_edges + 24 * tb.work_items + 16 * ts.work_edges + 24 * ts.work_items
    n_r = int((db != inf).sum())
    min_bytes = (8 + 16) * mr + (24 + 24) * n_r
    achieved = alg_bytes / (t_step_ms / 1e3) / 1e9
    achieved_min = min_bytes / (t_step_ms / 1e3) / 1e9

    out = {
        "metric": ("GTEPS on 1 MI355X (batched+learned vs plain, bit-exact); us per barrier"
                   if mode != gdb.SEL_NONE else
                   "GTEPS on 1 MI355X (batched vs plain, bit-exact; learned selector off); us per barrier"),
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.median(walls)) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"Lonestar worklist BFS + SSSP, {GRID_SIDE}x{GRID_SIDE} 4-neighbour grid, "
                        "5% edges removed (seed 2), weights U[1,255], source 0",
            "n": g.n, "m": g.m, "m_reached": mr,
            "engine": ("batched, owner-partitioned ranges (BFS and delta-band SSSP teams in one persistent "
                       "launch; each group holds a contiguous vertex range's distances in LDS)" if owned else
                       "batched (BFS and near/far SSSP teams in one persistent launch)"),
            "owned": owned,
            "wgs": wgs, "groups": groups, "team_groups": team, "delta": delta,
            "local_hops": hops, "lds_queue": qcap, "shape_source": "tuning table" if not (args.shape or args.tune)
            else ("--shape" if args.shape else "--tune"),
            "selector": {gdb.SEL_NONE: "none (push on the whole team)", gdb.SEL_MLP: "learned MLP"}[mode],
            "narrow_groups": narrow,
            "arm_iterations": {"bfs": list(last_tasks[0].arm_iters), "sssp": list(last_tasks[1].arm_iters)},
            "participating": bst.participating, "bfs_iterations": rounds[0], "sssp_iterations": rounds[1],
        },
        "task_ms": {"bfs": round(last_tasks[0].time_us / 1e3, 3), "sssp": round(last_tasks[1].time_us / 1e3, 3)},
        "kernel_ms": round(t_step_ms, 3),
        "tuned_single": {"gteps": round(2 * mr / (tuned_ms / 1e3) / 1e9, 4), "ms": round(tuned_ms, 3)},
        "plain": {"gteps": round(plain_gteps, 4), "ms": round(plain_ms, 3),
                  "participating": pss.participating, "barriers": psb.barriers + pss.barriers},
        "speedup_vs_plain": round(gteps / plain_gteps, 3),
        "bit_exact_vs_plain": bit_exact,
        "verify_errors": int(nerr),
        "barrier": {"kind_at_team_shape": "counter" if P < 64 else "tree",
                    "us_per_barrier": round(us_bar, 3), "us_per_relaunch": round(us_relaunch, 3),
                    "us_per_relaunch_with_readback": round(us_relaunch_sync, 3),
                    "ratio": round(us_bar / us_relaunch, 3), "P": P, "wgs": bwgs, "table": bars},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic_from_profiles((wgs, groups, delta, hops, mode, int(owned))),
            "kernel": "k_batch<WGS> (BFS + SSSP teams)",
            "algorithmic_bytes": int(alg_bytes),
            "minimal_bytes": int(min_bytes),
            "achieved_minimal": round(achieved_min, 2),
            "frac_minimal": round(achieved_min / HBM_PEAK_GBS, 5),
        },
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(g, db, ds)
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0 if bit_exact and nerr == 0 else 3


if __name__ == "__main__":
    sys.exit(main())
