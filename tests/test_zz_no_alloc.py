"""gdb_sssp_worklist_noalloc (no device allocation: its queues live in the
library's static per-GPU scratch array) against gdb_sssp_worklist on the
same inputs, bit for bit, through the C ABI: two calls in a row, the second
on a smaller graph (the array keeps the first call's contents), for the
tuned and plain persistent schedules and the relaunch schedule; and the
size limit is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle

pytestmark = pytest.mark.gpu

SCHEDS = [dict(variant=gdb.VARIANT_TUNED, wgs=256, groups=64),
          dict(variant=gdb.VARIANT_PLAIN, wgs=256, groups=32),
          dict(schedule=gdb.SCHED_RELAUNCH, wgs=256)]


def solve(fn, ctx, dg, out, opts):
    st = gdb.Stats()
    o = gdb.Opts.make(**opts)
    gdb.check(fn(ctx.h, dg.n, dg.m, dg.row.ptr, dg.col.ptr, dg.w.ptr, 0, out.ptr, C.byref(o), C.byref(st)),
              fn.__name__)
    return out.to_numpy(), st


@pytest.mark.parametrize("opts", SCHEDS, ids=["tuned", "plain", "relaunch"])
def test_noalloc_matches_allocating_entry_point(opts):
    L = gdb.lib()
    graphs = [gdb.gen_grid(600, 600, 0.05, seed=2, wlo=1, whi=255),
              gdb.gen_rmat(12, 8, seed=5, wlo=1, whi=100)]  # second call: smaller input
    with gdb.Context() as ctx:
        for g in graphs:
            dg = gdb.DeviceCSR(g)
            a_out = gdb.DeviceBuffer.empty(g.n, np.uint32)
            b_out = gdb.DeviceBuffer.from_numpy(np.full(g.n, 12345, np.uint32))
            print(f"noalloc vs allocating sssp: n={g.n} m={g.m} {opts}", flush=True)
            want, sa = solve(L.gdb_sssp_worklist, ctx, dg, a_out, opts)
            got, sb = solve(L.gdb_sssp_worklist_noalloc, ctx, dg, b_out, opts)
            assert (got == want).all()
            assert (got.astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all()
            assert sb.participating >= 1 and sa.participating >= 1


def test_noalloc_refuses_inputs_above_the_limit():
    g = gdb.gen_grid(8, 8, 0.0, seed=1, wlo=1, whi=9)
    with gdb.Context() as ctx:
        dg = gdb.DeviceCSR(g)
        out = gdb.DeviceBuffer.empty(g.n, np.uint32)
        st, o = gdb.Stats(), gdb.Opts.make()
        rc = gdb.lib().gdb_sssp_worklist_noalloc(ctx.h, (1 << 22) + 1, dg.m, dg.row.ptr, dg.col.ptr, dg.w.ptr, 0,
                                                 out.ptr, C.byref(o), C.byref(st))
        assert rc == -7  # GDB_E_SCRATCH
        # The context is still usable afterwards.
        got, _ = solve(gdb.lib().gdb_sssp_worklist_noalloc, ctx, dg, out, SCHEDS[0])
        assert (got.astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all()
