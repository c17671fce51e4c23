"""Headline benchmark: GTEPS of the worklist BFS + SSSP solve on a 2048x2048
4-neighbour grid (5% of edges removed, seed 2, weights U[1,255], source 0),
on one MI355X.

A step is one BFS solve plus one SSSP solve with the graph resident in HBM.
value = GTEPS of the tuned persistent engine = 2 * m_r / (BFS + SSSP kernel
time), with m_r = sum of out-degrees of reached vertices.  The plain
persistent engine (the reference kernels' structure, default launch) runs on
the same input; outputs of both engines are compared bit for bit and checked
with dverifysolution.  Also reported: us per barrier vs us per relaunch at the
tuned launch shape, the roofline of the dominant kernel, and a CPU baseline
(the oracle's single-thread restatement on the same graph).

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 2]
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def reached_edges(g, dist, inf) -> int:
    deg = np.diff(g.row).astype(np.int64)
    return int(deg[dist != inf].sum())


def traffic_from_profiles(kernel_prefix: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if any."""
    path = ROOT / "profiles" / "wl_tuned_c2_traffic.json"
    try:
        d = json.loads(path.read_text())
        return d.get(kernel_prefix)
    except Exception:
        return None


def cpu_baseline(g, seconds_cap: float = 30.0) -> dict:
    import oracle

    t0 = time.perf_counter()
    dist_b, _ = oracle.bfs_worklist_ref(g.row, g.col, 0)
    t1 = time.perf_counter()
    dist_s, _ = oracle.sssp_worklist_ref(g.row, g.col, g.w, 0)
    t2 = time.perf_counter()
    mr = reached_edges(g, dist_b, oracle.LONESTAR_INF) + reached_edges(g, dist_s, oracle.LONESTAR_INF)
    return {
        "value": mr / (t2 - t0) / 1e9,
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"one BFS + one SSSP worklist solve of the full {GRID_SIDE}x{GRID_SIDE} grid "
                  f"(oracle restatement, single thread): bfs {t1 - t0:.2f}s, sssp {t2 - t1:.2f}s",
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-tune", action="store_true", help="use the default launch shape")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build runs a single MI355X)", file=sys.stderr)
        return 2

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

    def solve_batched(wgs, groups, delta, hops=0, mode=gdb.SEL_NONE):
        # One launch, two teams: BFS and near/far SSSP side by side.  The
        # grid is symmetric, so it is its own in-edge CSR for pull arms.
        tasks = [gdb.make_task(gdb.TASK_BFS, dg, out_b, 0, hops=hops, transpose=dg),
                 gdb.make_task(gdb.TASK_SSSP, dg, out_s, 0, delta=delta, hops=hops, transpose=dg)]
        rounds, st = gdb.batch_solve(ctx, tasks, max(1, groups // 2),
                                     gdb.Opts.make(wgs=wgs, groups=groups), selector=selector(mode))
        last_tasks[:] = tasks
        return rounds, st

    # Autotuning of the batched engine (untimed): launch shape, near/far
    # band, local hops, and whether the learned selector schedules the
    # iterations (it needs hops = 0).
    modes = [gdb.SEL_NONE] + ([gdb.SEL_MLP] if weights is not None else [])
    if args.no_tune:
        wgs, groups, delta, hops, mode = 256, 64, 4096, 0, modes[-1]
    else:
        trials = []
        for w in (256, 512, 1024):
            for gr in (16, 32, 64, 128):
                gr = min(gr, gdb.max_resident_groups(w, 0))
                for dl in (2048, 4096, 8192):
                    for hp in (0, 4, 16):
                        for md in modes:
                            if md != gdb.SEL_NONE and hp > 0:
                                continue
                            _, st = solve_batched(w, gr, dl, hp, md)
                            trials.append((st.kernel_ms, w, gr, dl, hp, md))
        trials.sort()
        _, wgs, groups, delta, hops, mode = trials[0]
        log(f"autotune: wgs={wgs} groups={groups} delta={delta} hops={hops} selector={mode} "
            f"({trials[0][0]:.2f} ms); {len(trials)} shapes tried")

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

    # Barrier vs relaunch at the team launch shape (one team = groups / 2).
    P = max(1, groups // 2)
    K = 4000
    t0_ms, _ = ctx.barrier_bench(wgs, P, 0)
    tk_ms, pb = ctx.barrier_bench(wgs, P, K)
    us_bar = (tk_ms - t0_ms) * 1e3 / K
    us_relaunch = ctx.relaunch_bench(wgs, P, K, 0) * 1e3 / K
    us_relaunch_sync = ctx.relaunch_bench(wgs, P, 500, 1) * 1e3 / 500

    # Roofline of the dominant kernel (the batched kernel): algorithmic bytes
    # = BFS 8 B per processed edge + 24 B per processed vertex, SSSP 16 B per
    # processed edge + 24 B per processed worklist item.
    n_items, n_edges = bst.work_items, bst.work_edges
    alg_bytes = 16 * n_edges + 24 * n_items
    achieved = alg_bytes / (t_step_ms / 1e3) / 1e9

    out = {
        "metric": "GTEPS on 1 MI355X (batched+learned vs plain, bit-exact); us per XF barrier",
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
            "engine": "batched (BFS and near/far SSSP teams in one persistent launch)",
            "wgs": wgs, "groups": groups, "team_groups": max(1, groups // 2), "delta": delta,
            "local_hops": hops,
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
        "barrier": {"us_per_barrier": round(us_bar, 3), "us_per_relaunch": round(us_relaunch, 3),
                    "us_per_relaunch_with_readback": round(us_relaunch_sync, 3),
                    "ratio": round(us_bar / us_relaunch, 3), "P": pb, "wgs": wgs},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic_from_profiles("batch"),
            "kernel": "k_batch<WGS> (BFS + SSSP teams)",
            "algorithmic_bytes": int(alg_bytes),
        },
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(g)
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0 if bit_exact and nerr == 0 else 3


if __name__ == "__main__":
    sys.exit(main())
