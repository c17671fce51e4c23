"""Occupancy and protocol-timing suite: the counterpart of the reference's
occupancy_tests/scripts/run_occupancy_tests.py (true occupancy, then
discovered occupancy per lock kind over repeated runs: list, mean, standard
deviation) and run_occupancy_timing.py (protocol time per lock kind), for
the four configurations {min, max} local memory x {max, min} workgroup size
(run_occupancy_tests.py:117-129, 176-188).

Differences from the reference scripts, by design:
  - true occupancy comes from the occupancy API and is then checked with a
    skip-mode launch of exactly that many groups (every group takes part, one
    barrier must complete); the reference finds it by binary search with
    skip-mode launches that deadlock above it and a 7 s timeout
    (run_occupancy_tests.py:53-68), which a shared GPU must not do, so a
    launch one group above the bound is shown to be refused instead;
  - the lock-free protocol is measured beside the ticket and spin locks.

Usage: python -m gpu_discovery_barrier_amd.occupancy_suite [--iterations 10]
           [--groups 1000] [--json PATH]
"""
from __future__ import annotations

import argparse
import json
import math
import re
import sys

from . import api as gdb

LOCKS = (("spinlock", gdb.PROTOCOL_SPIN), ("ticket", gdb.PROTOCOL_TICKET), ("lock-free", gdb.PROTOCOL_LOCKFREE))


def avg(xs):
    return sum(xs) / float(len(xs))


def std_dev(xs):
    """Population standard deviation (run_occupancy_tests.py:200-205)."""
    m = avg(xs)
    return math.sqrt(avg([(x - m) ** 2 for x in xs]))


def device_info() -> dict:
    text = gdb.device_query()

    def field(name, num=True):
        m = re.search(rf"{name}: (.*)", text)
        if not m:
            return None
        return int(re.findall(r"\d+", m.group(1))[0]) if num else m.group(1).strip()

    return {"name": field("device name", False), "arch": field("arch", False),
            "local_mem": field("LDS per workgroup"), "max_wgs": field("max workgroup size"),
            "compute_units": field("compute units")}


def configs(info: dict):
    """(label, wgs, lms): max/min local memory x max/min workgroup size."""
    out = []
    for loc in ("max local memory", "min local memory"):
        for wg in ("max workgroup size", "min workgroup size"):
            wgs = info["max_wgs"] if wg.startswith("max") else 1
            lms = info["local_mem"] - 128 if loc.startswith("max") else 1
            out.append((f"{loc}\n{wg}", wgs, lms))
    return out


MAX_GROUPS = 8192  # barrier flag slots of the device context


def true_occupancy(ctx: gdb.Context, wgs: int, lms: int) -> dict:
    # The kernel itself declares 4 bytes of LDS next to the requested ones.
    bound = gdb.max_resident_groups(wgs, lms + 4)
    checked = min(bound, MAX_GROUPS)
    ok = ctx.occupancy_test(checked, wgs, lms, prot=False) == checked
    try:
        ctx.occupancy_test(bound + 1, wgs, lms, prot=False)
        refused = False
    except gdb.GdbError as e:
        refused = e.code in (gdb.GDB_E_UNSAFE, gdb.GDB_E_INVALID)
    return {"occupancy": bound, "skip_mode_checked_at": checked, "skip_mode_at_bound_completes": ok,
            "one_more_refused": refused}


def run(iterations: int, groups: int) -> dict:
    info = device_info()
    ctx = gdb.Context()
    data = []
    for label, wgs, lms in configs(info):
        d = {"config": label, "wgs": wgs, "lms": lms, **true_occupancy(ctx, wgs, lms)}
        for name, kind in LOCKS:
            ctx.configure(kind, gdb.BARRIER_AUTO)
            occ = [ctx.occupancy_test(groups, wgs, lms, prot=True) for _ in range(iterations)]
            ns = [ctx.time_protocol(groups, wgs, lms) for _ in range(iterations)]
            d[name] = {"occupancy_list": occ, "average": avg(occ), "std_dev": std_dev(occ),
                       "protocol_ms_average": avg(ns) / 1e6}
        ctx.configure(gdb.PROTOCOL_LOCKFREE, gdb.BARRIER_AUTO)
        data.append(d)
    ctx.close()
    return {"device": info, "groups_launched": groups, "iterations": iterations, "configs": data}


def report(res: dict) -> str:
    """The reference's text layout (mk_header, pp_data)."""
    info = res["device"]
    lines = [f"chip: {info['name']} ({info['arch']})", f"local memory:{info['local_mem']}",
             f"max workgroup size:{info['max_wgs']}", f"number of multiprocessors:{info['compute_units']}",
             "-----------------------------------"]
    for d in res["configs"]:
        lines += [d["config"], "---", f"true occupancy: {d['occupancy']}"]
        for name, _ in LOCKS:
            r = d[name]
            lines += [f"{name} occupancy list: {r['occupancy_list']}",
                      f"{name} average occupancy: {r['average']}",
                      f"{name} occupancy standard deviation: {r['std_dev']}",
                      f"{name} protocol time (ms, average): {r['protocol_ms_average']:.4f}"]
        lines.append("-----------------------------------")
    return "\n".join(lines)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--groups", type=int, default=1000, help="groups launched per run (the reference's 1000)")
    ap.add_argument("--json", default="")
    args = ap.parse_args()
    res = run(args.iterations, args.groups)
    print(report(res))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
