"""Trains the per-iteration schedule selector on traces the engine logs
itself, and evaluates it against the plain schedule on unseen graphs.

  collect  runs BFS and SSSP tasks with uniformly random arms (GDB_SEL_RANDOM)
           on the training graphs and keeps the trace records
  train    Adam on the device (gdb_mlp_train) over those records
  eval     kernel time of push-only (GDB_SEL_NONE), the heuristic and the
           MLP on unseen graphs; outputs are compared bit for bit with the
           push-only run

Training graphs: R-MAT scale 14 (seed 1), the 2048^2 grid (seed 2) and
R-MAT scale 22 (seed 3), several sources each.  Evaluation graphs: R-MAT
scales 20-23 and grids 1024-4096 with other seeds.

Usage: python -m gpu_discovery_barrier_amd.selector_train [--quick]
           [--weights PATH] [--report PATH] [--traces PATH]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import api as gdb

PKG = Path(__file__).resolve().parent
WEIGHTS = PKG / "selector_weights.npy"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Graph:
    """A symmetric graph resident on the device (it is its own transpose)."""

    def __init__(self, name: str, g: gdb.CSR):
        self.name, self.g = name, g
        self.dg = gdb.DeviceCSR(g)
        self.out = gdb.DeviceBuffer.empty(g.n, np.uint32)

    def task(self, algo: int, source: int, delta: int) -> gdb.Task:
        return gdb.make_task(algo, self.dg, self.out, source, delta=delta if algo == gdb.TASK_SSSP else 0,
                             transpose=self.dg)


def make(name: str) -> Graph:
    kind, a, seed = name.split(":")
    t0 = time.time()
    if kind == "rmat":
        g = gdb.gen_rmat(int(a), 16, seed=int(seed), wlo=1, whi=255)
    else:
        s = int(a)
        g = gdb.gen_grid(s, s, 0.05, seed=int(seed), wlo=1, whi=255)
    log(f"  {name}: n={g.n} m={g.m} ({time.time() - t0:.1f}s)")
    return Graph(name, g)


def launch_shape(g: gdb.CSR) -> tuple[int, int]:
    """(wgs, groups) of a single-task launch: one team over every group."""
    return (256, 256) if g.m > 50_000_000 else (256, 128)


def delta_of(g: gdb.CSR) -> int:
    return 4096 if g.m > 1_000_000 else 512


def solve(ctx, gr: Graph, algo: int, source: int, sel: gdb.Selector | None):
    wgs, groups = launch_shape(gr.g)
    t = gr.task(algo, source, delta_of(gr.g))
    _, st = gdb.batch_solve(ctx, [t], groups, gdb.Opts.make(wgs=wgs, groups=groups), selector=sel)
    return t, st


NARROW = 2


def steady_records(rec: np.ndarray) -> np.ndarray:
    """Records with a measured cost whose iteration kept the arm of the
    iteration before (same task, consecutive iterations): their cost holds
    no wake-up of parked groups or other switching cost.  Records of one
    solve are contiguous and in iteration order."""
    rec = np.asarray(rec)
    if len(rec) < 2:
        return rec[:0]
    prev, cur = rec[:-1], rec[1:]
    keep = ((cur[:, 17] > 0) & (prev[:, 18] == cur[:, 18]) & (prev[:, 19] == cur[:, 19] - 1)
            & (prev[:, 16] == cur[:, 16]))
    return np.ascontiguousarray(cur[keep])


def balanced(parts: list[np.ndarray], max_factor: int = 64) -> np.ndarray:
    """Concatenates per-graph record sets, repeating the smaller ones (up to
    max_factor times) so that each graph weighs about the same in training:
    a high-diameter grid logs thousands of iterations, an R-MAT graph ten."""
    parts = [p for p in parts if len(p)]
    if not parts:
        return np.zeros((0, gdb.TRACE_STRIDE), np.float32)
    big = max(len(p) for p in parts)
    out = [np.tile(p, (min(max_factor, max(1, big // len(p))), 1)) for p in parts]
    return np.ascontiguousarray(np.concatenate(out))


def collect(ctx, names, sources_per_graph, cap, seed):
    """Random-arm runs on each graph; returns the raw records of each graph."""
    trace = gdb.DeviceBuffer.from_numpy(np.zeros(cap * gdb.TRACE_STRIDE, np.float32))
    cnt = gdb.DeviceBuffer.from_numpy(np.zeros(1, np.int32))
    runs = 0
    bounds = [0]
    for name in names:
        gr = make(name)
        rng = np.random.default_rng(seed)
        reached = np.flatnonzero(np.diff(gr.g.row) > 0)
        srcs = [0] + [int(x) for x in rng.choice(reached, sources_per_graph - 1, replace=False)]
        for si, s in enumerate(srcs):
            for algo in (gdb.TASK_BFS, gdb.TASK_SSSP):
                sel = gdb.make_selector(gdb.SEL_RANDOM, narrow_groups=NARROW, seed=seed * 7919 + runs,
                                        trace=trace, trace_count=cnt)
                t, st = solve(ctx, gr, algo, s, sel)
                runs += 1
        log(f"  traces so far: {int(cnt.to_numpy()[0])}")
        bounds.append(min(int(cnt.to_numpy()[0]), cap))
        del gr
    rec = trace.to_numpy().reshape(cap, gdb.TRACE_STRIDE)
    return [rec[bounds[i]:bounds[i + 1]].copy() for i in range(len(names))]


def evaluate(ctx, names, weights: gdb.DeviceBuffer, reps: int, fixed_arms: bool = False):
    rows = []
    for name in names:
        gr = make(name)
        for algo in (gdb.TASK_BFS, gdb.TASK_SSSP):
            res = {}
            ref = None
            cands = [("push", None),
                     ("heuristic", gdb.make_selector(gdb.SEL_HEURISTIC, narrow_groups=NARROW)),
                     ("mlp", gdb.make_selector(gdb.SEL_MLP, narrow_groups=NARROW, weights=weights))]
            if fixed_arms:
                cands += [(f"arm{a}", gdb.make_selector(gdb.SEL_FIXED, fixed_arm=a, narrow_groups=NARROW))
                          for a in range(1, 4)]
            for label, sel in cands:
                times = []
                for _ in range(reps):
                    t, st = solve(ctx, gr, algo, 0, sel)
                    times.append(st.kernel_ms)
                out = gr.out.to_numpy()
                if ref is None:
                    ref = out
                res[label] = {"ms": float(np.median(times)), "arms": list(t.arm_iters),
                              "bit_exact": bool((out == ref).all())}
            row = {"graph": name, "algo": "bfs" if algo == gdb.TASK_BFS else "sssp", **res,
                   "mlp_speedup": res["push"]["ms"] / res["mlp"]["ms"],
                   "heuristic_speedup": res["push"]["ms"] / res["heuristic"]["ms"]}
            log(f"  {name} {row['algo']}: " + ", ".join(f"{k} {v['ms']:.2f} ms {v['arms']}"
                                                     + ("" if v["bit_exact"] else " NOT BIT-EXACT")
                                                     for k, v in res.items()))
            rows.append(row)
        del gr
    return rows


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--quick", action="store_true", help="small graphs only (a smoke run)")
    ap.add_argument("--weights", default=str(WEIGHTS))
    ap.add_argument("--report", default="")
    ap.add_argument("--traces", default="")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--arms", action="store_true", help="also time each arm alone")
    args = ap.parse_args()
    if args.quick:
        train = ["rmat:12:1", "grid:256:2", "rmat:16:3"]
        test = ["rmat:15:51", "grid:384:7"]
        nsrc, reps = 2, 1
    else:
        train = ["rmat:14:1", "grid:2048:2", "rmat:22:3"]
        test = ["rmat:20:50", "rmat:21:51", "rmat:22:52", "rmat:23:53", "grid:1024:7", "grid:2048:8",
                "grid:4096:9"]
        nsrc, reps = 6, 3
    ctx = gdb.Context()
    report = {"train_graphs": train, "test_graphs": test, "narrow_groups": NARROW}
    if not args.eval_only:
        log("collecting traces ...")
        raw = collect(ctx, train, nsrc, 400_000, seed=5)
        if args.traces:
            np.save(args.traces, np.concatenate(raw))
        steady = [steady_records(r) for r in raw]
        rec = balanced(steady)
        log(f"{sum(len(r) for r in raw)} records, steady per graph {[len(r) for r in steady]}, "
            f"{len(rec)} after balancing; arms {np.bincount(rec[:, 16].astype(int), minlength=4).tolist()}")
        drec = gdb.DeviceBuffer.from_numpy(rec)
        p0 = gdb.mlp_init(5)
        _, l0, _ = gdb.mlp_grad(ctx, gdb.DeviceBuffer.from_numpy(gdb.bf16_round(p0)), drec, len(rec))
        t0 = time.time()
        p, w, l1 = gdb.mlp_train(ctx, p0, drec, len(rec), steps=args.steps, batch=512, lr=2e-3, seed=5)
        log(f"trained: loss {l0:.3f} -> {l1:.4f} in {time.time() - t0:.1f}s")
        np.save(args.weights, gdb.bf16_round(p))
        report.update({"records": int(sum(len(r) for r in raw)), "steady_records": [len(r) for r in steady],
                       "loss_initial": l0, "loss_final": l1, "steps": args.steps})
    wb = np.load(args.weights)
    weights = gdb.DeviceBuffer.from_numpy(wb.astype(np.uint16))
    log("evaluating ...")
    report["eval"] = evaluate(ctx, test, weights, reps, args.arms)
    ctx.close()
    if args.report:
        Path(args.report).write_text(json.dumps(report, indent=1))
    print(json.dumps({"eval": [(r["graph"], r["algo"], round(r["mlp_speedup"], 3),
                                round(r["heuristic_speedup"], 3)) for r in report["eval"]]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
