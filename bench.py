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

    # Launch-shape autotuning of the tuned engine (untimed).
    if args.no_tune:
        wgs, groups = 256, 0
    else:
        def run(w, gr):
            _, sb, _, ss = solve(gdb.VARIANT_TUNED, w, gr)
            return sb.kernel_ms + ss.kernel_ms
        best = autotune.tune(f"wl_bfs_sssp:grid{GRID_SIDE}:{g.m}", run, repeats=1, use_cache=False)
        wgs, groups = best["wgs"], best["groups"]
        log(f"autotune: wgs={wgs} groups={groups} ({best['ms']:.2f} ms); "
            f"{len(best['trials'])} shapes tried")

    for _ in range(args.warmup):
        solve(gdb.VARIANT_TUNED, wgs, groups)
    bfs_ms, sssp_ms, walls = [], [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        db, sb, ds, ss = solve(gdb.VARIANT_TUNED, wgs, groups)
        walls.append(time.perf_counter() - t0)
        bfs_ms.append(sb.kernel_ms)
        sssp_ms.append(ss.kernel_ms)

    inf = gdb.LONESTAR_INF
    mr = reached_edges(g, db, inf)
    t_step_ms = float(np.median(np.array(bfs_ms) + np.array(sssp_ms)))
    gteps = 2 * mr / (t_step_ms / 1e3) / 1e9

    # Plain engine (reference kernel structure, default launch shape).
    pdb, psb, pds, pss = solve(gdb.VARIANT_PLAIN, 256, 0)
    plain_ms = psb.kernel_ms + pss.kernel_ms
    plain_gteps = 2 * mr / (plain_ms / 1e3) / 1e9
    bit_exact = bool((pdb == db).all() and (pds == ds).all())
    nerr = gdb.verify_distances(ctx, dg, gdb.DeviceBuffer.from_numpy(ds), weighted=True)
    nerr += gdb.verify_distances(ctx, dg, gdb.DeviceBuffer.from_numpy(db), weighted=False)

    # Barrier vs relaunch at the tuned launch shape.
    P = ss.participating
    K = 4000
    t0_ms, _ = ctx.barrier_bench(wgs, P, 0)
    tk_ms, pb = ctx.barrier_bench(wgs, P, K)
    us_bar = (tk_ms - t0_ms) * 1e3 / K
    us_relaunch = ctx.relaunch_bench(wgs, P, K, 0) * 1e3 / K
    us_relaunch_sync = ctx.relaunch_bench(wgs, P, 500, 1) * 1e3 / 500

    # Roofline of the dominant kernel (the SSSP worklist kernel): algorithmic
    # bytes = 16 B per processed edge (dst, weight, dist[src], dist[dst]) +
    # 24 B per processed worklist item.
    sssp_bytes = 16 * ss.work_edges + 24 * ss.work_items
    sssp_kernel_ms = float(np.median(sssp_ms))
    achieved = sssp_bytes / (sssp_kernel_ms / 1e3) / 1e9

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
            "engine": "tuned persistent", "wgs": wgs, "groups": groups, "participating": P,
            "bfs_levels": sb.rounds, "sssp_iterations": ss.rounds,
        },
        "kernel_ms": {"bfs": round(float(np.median(bfs_ms)), 3), "sssp": round(sssp_kernel_ms, 3)},
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
            "traffic": traffic_from_profiles("sssp"),
            "kernel": "k_wl_tuned<WGS,true> (worklist SSSP)",
            "algorithmic_bytes": int(sssp_bytes),
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
