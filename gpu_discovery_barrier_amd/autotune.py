"""Launch-shape autotuner and the tuning table the library reads.

Counterpart of the reference's work-group-size tuner
(pannotia/scripts/tune_wg_size.py:106-124, lonestar_gpu_opencl/scripts/
tune_wg_size.py), which times every work-group size per app and graph and
tabulates the fastest (wg_size_data.py:7-79) for the suite runner to use.
Here the search space also covers the number of launched groups (a barrier's
cost grows with the groups it joins), the team size, near/far band and
local-hop budget of the batched engine, and its LDS queue tile; the winners
go into a text table (tune_table.txt beside libgdb_amd.so, format in
include/gdb_amd.h, gdb_tune_load) that every solver consults when it is
called with wgs = 0, so the library, the CLIs (WORKGROUP_SIZE 0) and
bench.py all run the tuned shapes.

Usage: python -m gpu_discovery_barrier_amd.autotune [--quick] [--out PATH]
          [--only color,mis,bc,batch_grid,batch_many,worklist]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import api

TABLE = api.PKG / "tune_table.txt"
COLUMNS = ("solver", "n_max", "wgs", "groups", "team_groups", "delta", "hops", "lds_queue")
# Dynamic LDS of a batch with an owned task: a CU's 160 KiB less the kernel's
# static LDS, so one such group fits per CU.
OWNED_LDS = 163840 - 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- the table
def read_table(path: Path | str = TABLE) -> list[dict]:
    rows = []
    for line in Path(path).read_text().splitlines():
        line = line.split("#", 1)[0].split()
        if not line:
            continue
        if len(line) != len(COLUMNS):
            raise ValueError(f"malformed tuning row: {line}")
        rows.append({k: (v if k == "solver" else int(v)) for k, v in zip(COLUMNS, line)})
    return rows


def write_table(rows: list[dict], path: Path | str = TABLE, header: str = "") -> None:
    lines = ["# Tuning table of libgdb_amd.so: solver n_max wgs groups team_groups delta hops lds_queue",
             "# (written by python -m gpu_discovery_barrier_amd.autotune; a solver called with wgs = 0",
             "#  uses the row for its solver with the smallest n_max >= n)"]
    lines += [f"# {h}" for h in header.splitlines() if h]
    for r in sorted(rows, key=lambda r: (r["solver"], r["n_max"])):
        lines.append(" ".join(str(r[k]) for k in COLUMNS))
    Path(path).write_text("\n".join(lines) + "\n")


def lookup(solver: str, n: int, path: Path | str = TABLE) -> dict | None:
    """The row the library would use (read from the file, no GPU needed)."""
    best = None
    for r in read_table(path):
        if r["solver"] == solver and r["n_max"] >= n and (best is None or r["n_max"] < best["n_max"]):
            best = r
    return best


def row(solver, n_max, wgs, groups, team_groups=0, delta=0, hops=0, lds_queue=0) -> dict:
    return dict(solver=solver, n_max=int(n_max), wgs=int(wgs), groups=int(groups), team_groups=int(team_groups),
                delta=int(delta), hops=int(hops), lds_queue=int(lds_queue))


# ---------------------------------------------------------------- the sweeps
def best_of(trials: list[tuple]) -> tuple:
    return min(trials, key=lambda t: t[0])


def shapes(wgs_list, groups_list):
    for w in wgs_list:
        cap = api.max_resident_groups(w, 0)
        for g in sorted({min(g, cap) for g in groups_list}):
            yield w, g


def tune_solver(ctx, name, graph_n, run, wgs_list, groups_list, repeats=2) -> tuple[dict, list]:
    trials = []
    for w, g in shapes(wgs_list, groups_list):
        ms = min(run(w, g) for _ in range(repeats))
        trials.append((ms, w, g))
    ms, w, g = best_of(trials)
    log(f"{name}: wgs {w} groups {g} {ms:.3f} ms ({len(trials)} shapes)")
    return row(name, graph_n, w, g), trials


def tune_color(ctx, quick=False):
    g = api.gen_rmat(14, 16, seed=1)
    dg = api.DeviceCSR(g, weights=False)
    nv = api.DeviceBuffer.from_numpy((api.permutation(g.n, 1).astype(np.float64) / (g.n + 1)).astype(np.float32))
    out = api.DeviceBuffer.empty(g.n, np.int32)
    run = lambda w, gr: api.color(ctx, dg, nv, api.Opts.make(wgs=w, groups=gr, variant=api.VARIANT_TUNED),
                                  out=out)[1].kernel_ms
    return tune_solver(ctx, "color", 1 << 20, run, (64, 128, 256) if quick else (64, 128, 256, 512, 1024),
                       (32, 64, 128, 256, 512))


def tune_mis_bc(ctx, quick=False):
    scale = 20 if quick else 22
    g = api.gen_rmat(scale, 16, seed=3)
    dg = api.DeviceCSR(g, weights=False)
    nv = api.DeviceBuffer.from_numpy((api.permutation(g.n, 3).astype(np.float64) / (g.n + 1)).astype(np.float32))
    out = api.DeviceBuffer.empty(g.n, np.int32)
    mis = lambda w, gr: api.mis(ctx, dg, nv, api.Opts.make(wgs=w, groups=gr, variant=api.VARIANT_TUNED),
                                out=out)[1].kernel_ms
    r1, t1 = tune_solver(ctx, "mis", 1 << 30, mis, (256, 512, 1024), (512, 1024, 2048, 4096))
    deg = np.diff(g.row)
    src = np.sort(np.random.default_rng(3).choice(np.flatnonzero(deg > 0), 16, replace=False)).astype(np.int32)
    bout = api.DeviceBuffer.empty(g.n, np.float32)
    bc = lambda w, gr: api.bc(ctx, dg, dg, src, api.Opts.make(wgs=w, groups=gr, variant=api.VARIANT_TUNED),
                              out=bout)[1].kernel_ms
    r2, t2 = tune_solver(ctx, "bc", 1 << 30, bc, (256, 512), (1024, 2048), repeats=1)
    return [r1, r2], t1 + t2


def tune_worklist(ctx, quick=False):
    side = 1024 if quick else 2048
    g = api.gen_grid(side, side, 0.05, seed=2, wlo=1, whi=255)
    dg = api.DeviceCSR(g)
    out = api.DeviceBuffer.empty(g.n, np.uint32)
    rows, trials = [], []
    for name, fn in (("bfs_wl", api.bfs_worklist), ("sssp_wl", api.sssp_worklist)):
        run = lambda w, gr: fn(ctx, dg, 0, api.Opts.make(wgs=w, groups=gr, variant=api.VARIANT_TUNED),
                               out=out)[1].kernel_ms
        r, t = tune_solver(ctx, name, 1 << 30, run, (256, 512), (64, 128, 256), repeats=1)
        rows.append(r)
        trials += t
    return rows, trials


def grid_batch_run(ctx, dg, out_b, out_s, w, gr, team, delta, hops, qcap, owned=False):
    tasks = [api.make_task(api.TASK_BFS, dg, out_b, 0, hops=hops, transpose=dg, owned=owned),
             api.make_task(api.TASK_SSSP, dg, out_s, 0, delta=delta, hops=hops, transpose=dg, owned=owned)]
    _, st = api.batch_solve(ctx, tasks, team, api.Opts.make(wgs=w, groups=gr, lds_queue=qcap))
    return st.kernel_ms


def tune_batch_grid(ctx, quick=False):
    """BFS + SSSP of one large grid side by side (two teams of groups / 2)."""
    g = api.gen_grid(2048, 2048, 0.05, seed=2, wlo=1, whi=255)
    dg = api.DeviceCSR(g)
    ob, os_ = api.DeviceBuffer.empty(g.n, np.uint32), api.DeviceBuffer.empty(g.n, np.uint32)
    trials = []
    space_w = (128, 256) if quick else (128, 256, 512)
    space_g = (128, 256) if quick else (64, 128, 256, 512)
    space_d = (4096,) if quick else (2048, 4096, 8192)
    space_h = (256,) if quick else (64, 256, 1024)
    space_q = (1024, 2048, 4096)
    for w, gr in shapes(space_w, space_g):
        for dl in space_d:
            for hp in space_h:
                for q in space_q:
                    ms = min(grid_batch_run(ctx, dg, ob, os_, w, gr, max(1, gr // 2), dl, hp, q) for _ in range(2))
                    trials.append((ms, w, gr, max(1, gr // 2), dl, hp, q))
    ms, w, gr, team, dl, hp, q = best_of(trials)
    log(f"batch (grid 2048): wgs {w} groups {gr} team {team} delta {dl} hops {hp} lds_queue {q} {ms:.3f} ms "
        f"({len(trials)} shapes)")
    rows = [row("batch", 1 << 30, w, gr, team, dl, hp, q)]
    # Owner-partitioned ranges (gdb_task.owned): one group per CU (the range's
    # distances fill the LDS), two teams of half the groups; the queue tile
    # is capped inside the kernel.  Recorded as its own row when it beats the
    # shared engine, which makes it the headline engine for this workload.
    otrials = []
    for w2 in ((1024,) if quick else (512, 1024)):
        gr2 = api.max_resident_groups(w2, OWNED_LDS)  # one group per CU
        for dl in ((6144,) if quick else (4096, 6144, 8192, 12288)):
            ms2 = min(grid_batch_run(ctx, dg, ob, os_, w2, gr2, gr2 // 2, dl, 0, 2048, owned=True) for _ in range(2))
            otrials.append((ms2, w2, gr2, gr2 // 2, dl, 0, 2048))
    oms, ow, ogr, oteam, odl, _, oq = best_of(otrials)
    log(f"batch owned (grid 2048): wgs {ow} groups {ogr} team {oteam} delta {odl} {oms:.3f} ms")
    if oms < ms:
        rows.append(row("batch_owned", 1 << 30, ow, ogr, oteam, odl, 0, oq))
    return rows, trials + [t + ("owned",) for t in otrials]


def tune_batch_many(ctx, quick=False):
    """Many small graphs (G(n,p) n in [256, 65536], grids side 16..256), one
    group per task: BFS, SSSP and colour of each."""
    rng = np.random.default_rng(4)
    count = 256 if quick else 1024
    gs = []
    for i in range(count // 2):
        n = int(np.floor(np.exp(rng.uniform(np.log(256), np.log(65536)))))
        gs.append(api.gen_gnp(n, 8.0, seed=4 + i, wlo=1, whi=100))
    for i in range(count // 2):
        sd = int(rng.integers(16, 257))
        gs.append(api.gen_grid(sd, sd, 0.0, seed=4 + 2048 + i, wlo=1, whi=100))
    dgs = [api.DeviceCSR(g) for g in gs]
    nvs = [api.DeviceBuffer.from_numpy(api.pannotia_values(g.n)) for g in gs]
    outs = [[api.DeviceBuffer.empty(g.n, np.uint32), api.DeviceBuffer.empty(g.n, np.uint32),
             api.DeviceBuffer.empty(g.n, np.int32)] for g in gs]

    def run(w, gr, dl):
        ts = []
        for dg, nv, o in zip(dgs, nvs, outs):
            ts.append(api.make_task(api.TASK_BFS, dg, o[0], 0))
            ts.append(api.make_task(api.TASK_SSSP, dg, o[1], 0, delta=dl))
            ts.append(api.make_task(api.TASK_COLOR, dg, o[2], 0, node_value=nv))
        return api.batch_solve(ctx, ts, 1, api.Opts.make(wgs=w, groups=gr))[1].kernel_ms

    # No local hops here, so the batch takes no LDS queue (lds_queue 0).
    trials = []
    for w, gr in shapes((256, 512), (512, 1024, 2048)):
        for dl in (0, 200):
            trials.append((run(w, gr, dl), w, gr, dl))
    ms, w, gr, dl = best_of(trials)
    log(f"batch (many graphs): wgs {w} groups {gr} delta {dl} {ms:.3f} ms ({len(trials)} shapes)")
    return row("batch", 65536, w, gr, 1, dl, 0, 0), trials


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--quick", action="store_true", help="smaller search spaces and graphs")
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--only", default="color,mis,bc,worklist,batch_grid,batch_many")
    ap.add_argument("--report", default="", help="also write every trial as JSON here")
    a = ap.parse_args()
    only = set(a.only.split(","))
    ctx = api.Context()
    rows, report = [], {}
    t0 = time.time()
    try:
        old = {(r["solver"], r["n_max"]): r for r in read_table(a.out)} if Path(a.out).exists() else {}
    except ValueError:
        old = {}
    if "color" in only:
        r, t = tune_color(ctx, a.quick)
        rows.append(r)
        report["color"] = t
    if only & {"mis", "bc"}:
        rs, t = tune_mis_bc(ctx, a.quick)
        rows += rs
        report["mis_bc"] = t
    if "worklist" in only:
        rs, t = tune_worklist(ctx, a.quick)
        rows += rs
        report["worklist"] = t
    if "batch_grid" in only:
        rs, t = tune_batch_grid(ctx, a.quick)
        rows += rs
        report["batch_grid"] = t
        # An owned row from an earlier run stands only if this run kept one.
        old = {k: v for k, v in old.items() if k[0] != "batch_owned"}
    if "batch_many" in only:
        r, t = tune_batch_many(ctx, a.quick)
        rows.append(r)
        report["batch_many"] = t
    ctx.close()
    merged = dict(old)
    for r in rows:
        merged[(r["solver"], r["n_max"])] = r
    # The first device-query line with a value (the marketing name can be empty).
    dev = [ln.strip() for ln in api.device_query().splitlines() if ln.split(":", 1)[-1].strip()]
    write_table(list(merged.values()), a.out,
                header=f"device: {dev[0] if dev else '?'}; tuned in {time.time() - t0:.0f}s"
                       f"{' (quick)' if a.quick else ''}")
    if a.report:
        Path(a.report).write_text(json.dumps(report))
    print(json.dumps({"table": a.out, "rows": rows}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
