"""A caller's own persistent kernel, compiled outside libgdb_amd.so against
the public device header (include/gdb_discovery.hpp), runs the library's
discovery protocol and barrier on a library context: the context's device
pointer comes from gdb_ctx_device_ptr, the launch from gdb_launch_on_ctx,
the timing from gdb_time_protocol_kernel (the reference's time_protocol,
discovery.h:71-105, which times the caller's kernel).  Checked by the
barrier invariants: ids form 0..P-1, per-round counts are P*k, stores made
before a barrier are visible after it."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb

pytestmark = pytest.mark.gpu
EXT = Path(__file__).resolve().parent / "ext" / "libgdb_user_kernel.so"


@pytest.fixture(scope="module")
def kernel():
    if not EXT.exists():
        raise RuntimeError(f"{EXT} is missing: build it with `make -C tests/ext`")
    lib = C.CDLL(str(EXT))
    lib.user_barrier_kernel_symbol.restype = C.c_void_p
    return lib, lib.user_barrier_kernel_symbol()


def declare():
    L = gdb.lib()
    L.gdb_ctx_device_ptr.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    L.gdb_launch_on_ctx.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_size_t]
    L.gdb_time_protocol_kernel.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_int,
                                           C.c_size_t, C.POINTER(C.c_double)]
    return L


def kernel_args(dctx, iters, count, data, stride, bad, ids):
    vals = [C.c_void_p(dctx), C.c_int(iters), C.c_void_p(count.ptr.value), C.c_void_p(data.ptr.value),
            C.c_int(stride), C.c_void_p(bad.ptr.value), C.c_void_p(ids.ptr.value)]
    arr = (C.c_void_p * len(vals))(*[C.cast(C.pointer(v), C.c_void_p) for v in vals])
    return arr, vals


@pytest.mark.parametrize("bar", [gdb.BARRIER_XF, gdb.BARRIER_COUNTER, gdb.BARRIER_TREE])
@pytest.mark.parametrize("wgs,groups", [(64, 5), (256, 300), (1024, 128)])
def test_user_kernel_barrier_invariants(kernel, bar, wgs, groups):
    L = declare()
    groups = min(groups, gdb.max_resident_groups(wgs, 0))
    with gdb.Context(bar_kind=bar) as ctx:
        dctx = C.c_void_p()
        gdb.check(L.gdb_ctx_device_ptr(ctx.h, C.byref(dctx)), "gdb_ctx_device_ptr")
        stride = groups * wgs
        for rep, iters in enumerate((50, 7)):  # two launches after one init
            count = gdb.DeviceBuffer.from_numpy(np.zeros(4, np.int32))
            data = gdb.DeviceBuffer.from_numpy(np.full(2 * stride, -1, np.int32))
            bad = gdb.DeviceBuffer.from_numpy(np.zeros(2, np.int32))
            ids = gdb.DeviceBuffer.from_numpy(np.zeros(stride, np.int32))
            args, keep = kernel_args(dctx.value, iters, count, data, stride, bad, ids)
            print(f"user kernel: bar {bar} wgs {wgs} groups {groups} iters {iters}", flush=True)
            gdb.check(L.gdb_launch_on_ctx(ctx.h, kernel[1], args, groups, wgs, 0), "gdb_launch_on_ctx")
            P = ctx.participating()
            assert 1 <= P <= groups
            assert int(bad.to_numpy().view(np.int64)[0]) == 0
            marks = ids.to_numpy()
            assert (marks[:P] == 1).all() and (marks[P:] == 0).all()


def test_time_protocol_times_the_callers_kernel(kernel):
    L = declare()
    with gdb.Context() as ctx:
        dctx = C.c_void_p()
        gdb.check(L.gdb_ctx_device_ptr(ctx.h, C.byref(dctx)), "gdb_ctx_device_ptr")
        wgs, ub = 256, 1000
        stride = ub * wgs
        count = gdb.DeviceBuffer.from_numpy(np.zeros(4, np.int32))
        data = gdb.DeviceBuffer.from_numpy(np.full(2 * stride, -1, np.int32))
        bad = gdb.DeviceBuffer.from_numpy(np.zeros(2, np.int32))
        ids = gdb.DeviceBuffer.from_numpy(np.zeros(stride, np.int32))
        times = []
        for iters in (0, 2000):
            args, keep = kernel_args(dctx.value, iters, count, data, stride, bad, ids)
            ns = C.c_double(0)
            print(f"time_protocol on the user kernel, {ub} groups, {iters} barriers", flush=True)
            gdb.check(L.gdb_time_protocol_kernel(ctx.h, kernel[1], args, ub, wgs, 0, C.byref(ns)),
                      "gdb_time_protocol_kernel")
            times.append(ns.value)
            assert 1 <= ctx.participating() <= min(ub, gdb.max_resident_groups(wgs, 0))
        assert int(bad.to_numpy().view(np.int64)[0]) == 0
        # 2000 barriers take measurably longer than the protocol alone.
        assert times[1] > times[0] > 0
