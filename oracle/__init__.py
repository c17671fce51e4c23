"""Python access to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (libgdb_amd.so and
its Python binding) never imports this package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
LONESTAR_INF = 1000000000
PANNOTIA_INF = 99999999

_lib = None
P = C.c_void_p
I = C.c_int


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_color_ref": (I, [I, P, P, P, P]),
            "orc_color_tb": (I, [I, P, P, P, P]),
            "orc_color_maxmin_ref": (I, [I, P, P, P, P]),
            "orc_mis_ref": (I, [I, P, P, P, P]),
            "orc_mis_tb": (I, [I, P, P, P, P]),
            "orc_sssp_pull_ref": (I, [I, P, P, P, I, P]),
            "orc_dijkstra_tb": (I, [I, P, P, P, I, C.c_longlong, P]),
            "orc_bellman_ford_tb": (I, [I, P, P, P, I, C.c_longlong, P]),
            "orc_bfs_worklist_ref": (I, [I, P, P, I, P]),
            "orc_sssp_worklist_ref": (I, [I, P, P, P, I, P]),
            "orc_bfs_tb": (I, [I, P, P, I, P]),
            "orc_verify_distances_ref": (C.c_longlong, [I, P, P, P, P]),
            "orc_bc_brandes_tb": (I, [I, P, P, P, I, P]),
            "orc_bc_pairs_tb": (I, [I, P, P, P, I, P]),
            "orc_bc_brandes_mt": (I, [I, P, P, P, I, I, P]),
            "orc_worklist_mt": (I, [I, P, P, P, I, I, P, C.POINTER(I)]),
            "orc_bc_ref_f32": (I, [I, P, P, P, P, P, I, P]),
            "orc_block_scan_ref": (I, [I, P, P]),
            "orc_dp_model": (I, [I, I, I, I, I, I, I, C.POINTER(I), P, C.POINTER(C.c_longlong)]),
            "orc_mlp_forward": (None, [P, P, P]),
            "orc_mst_kruskal": (I, [I, P, P, P, C.POINTER(C.c_ulonglong), C.POINTER(I), C.POINTER(I)]),
            "orc_mst_prim_tb": (I, [I, P, P, P, C.POINTER(C.c_ulonglong), C.POINTER(I), C.POINTER(I)]),
            "orc_mlp_grad": (I, [P, P, I, P, C.POINTER(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(P)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def color_ref(row, col, nv):
    row, col, nv = _i32(row), _i32(col), np.ascontiguousarray(nv, np.float32)
    n = row.size - 1
    out = np.empty(n, np.int32)
    r = lib().orc_color_ref(n, _p(row), _p(col), _p(nv), _p(out))
    return out, r


def color_tb(row, col, nv):
    row, col, nv = _i32(row), _i32(col), np.ascontiguousarray(nv, np.float32)
    n = row.size - 1
    out = np.empty(n, np.int32)
    r = lib().orc_color_tb(n, _p(row), _p(col), _p(nv), _p(out))
    return out, r


def color_maxmin_ref(row, col, nv):
    row, col, nv = _i32(row), _i32(col), np.ascontiguousarray(nv, np.float32)
    n = row.size - 1
    out = np.empty(n, np.int32)
    r = lib().orc_color_maxmin_ref(n, _p(row), _p(col), _p(nv), _p(out))
    return out, r


def mis_ref(row, col, nv):
    row, col, nv = _i32(row), _i32(col), np.ascontiguousarray(nv, np.float32)
    n = row.size - 1
    out = np.empty(n, np.int32)
    r = lib().orc_mis_ref(n, _p(row), _p(col), _p(nv), _p(out))
    return out, r


def mis_tb(row, col, nv):
    row, col, nv = _i32(row), _i32(col), np.ascontiguousarray(nv, np.float32)
    n = row.size - 1
    out = np.empty(n, np.int32)
    lib().orc_mis_tb(n, _p(row), _p(col), _p(nv), _p(out))
    return out


def sssp_pull_ref(row_t, col_t, w_t, src=0):
    row_t, col_t, w_t = _i32(row_t), _i32(col_t), _i32(w_t)
    n = row_t.size - 1
    out = np.empty(n, np.int32)
    r = lib().orc_sssp_pull_ref(n, _p(row_t), _p(col_t), _p(w_t), src, _p(out))
    return out, r


def dijkstra_tb(row, col, w, src=0, inf=LONESTAR_INF):
    row, col = _i32(row), _i32(col)
    w = None if w is None else _i32(w)
    n = row.size - 1
    out = np.empty(n, np.int64)
    lib().orc_dijkstra_tb(n, _p(row), _p(col), None if w is None else _p(w), src, inf, _p(out))
    return out


def bellman_ford_tb(row, col, w, src=0, inf=LONESTAR_INF):
    row, col = _i32(row), _i32(col)
    w = None if w is None else _i32(w)
    n = row.size - 1
    out = np.empty(n, np.int64)
    lib().orc_bellman_ford_tb(n, _p(row), _p(col), None if w is None else _p(w), src, inf, _p(out))
    return out


def bfs_worklist_ref(row, col, src=0):
    row, col = _i32(row), _i32(col)
    n = row.size - 1
    out = np.empty(n, np.uint32)
    it = lib().orc_bfs_worklist_ref(n, _p(row), _p(col), src, _p(out))
    return out, it


def sssp_worklist_ref(row, col, w, src=0):
    row, col, w = _i32(row), _i32(col), _i32(w)
    n = row.size - 1
    out = np.empty(n, np.uint32)
    it = lib().orc_sssp_worklist_ref(n, _p(row), _p(col), _p(w), src, _p(out))
    return out, it


def worklist_mt(row, col, w, src=0, threads=None):
    """Worklist BFS (w None) or SSSP on `threads` host threads with a CPU XF
    barrier: (dist uint32[n], iterations)."""
    import os
    row, col = _i32(row), _i32(col)
    w = None if w is None else _i32(w)
    threads = threads or min(16, os.cpu_count() or 1)
    n = row.size - 1
    out = np.empty(n, np.uint32)
    it = C.c_int(0)
    rc = lib().orc_worklist_mt(n, _p(row), _p(col), None if w is None else _p(w), src, int(threads), _p(out),
                               C.byref(it))
    if rc:
        raise MemoryError("orc_worklist_mt")
    return out, it.value


def bfs_tb(row, col, src=0):
    row, col = _i32(row), _i32(col)
    n = row.size - 1
    out = np.empty(n, np.uint32)
    lib().orc_bfs_tb(n, _p(row), _p(col), src, _p(out))
    return out


def verify_distances_ref(row, col, w, dist):
    row, col = _i32(row), _i32(col)
    w = None if w is None else _i32(w)
    dist = np.ascontiguousarray(dist, np.uint32)
    n = row.size - 1
    return lib().orc_verify_distances_ref(n, _p(row), _p(col), None if w is None else _p(w), _p(dist))


def bc_brandes_tb(row, col, sources):
    row, col, sources = _i32(row), _i32(col), _i32(sources)
    n = row.size - 1
    out = np.empty(n, np.float64)
    lib().orc_bc_brandes_tb(n, _p(row), _p(col), _p(sources), sources.size, _p(out))
    return out


def bc_brandes_mt(row, col, sources, threads=None):
    """bc_brandes_tb with the sources spread over `threads` host threads."""
    import os
    row, col, sources = _i32(row), _i32(col), _i32(sources)
    threads = threads or min(16, os.cpu_count() or 1)
    n = row.size - 1
    out = np.empty(n, np.float64)
    rc = lib().orc_bc_brandes_mt(n, _p(row), _p(col), _p(sources), sources.size, int(threads), _p(out))
    if rc:
        raise MemoryError("orc_bc_brandes_mt")
    return out


def bc_pairs_tb(row, col, sources):
    row, col, sources = _i32(row), _i32(col), _i32(sources)
    n = row.size - 1
    out = np.empty(n, np.float64)
    lib().orc_bc_pairs_tb(n, _p(row), _p(col), _p(sources), sources.size, _p(out))
    return out


def bc_ref_f32(row, col, row_t, col_t, sources):
    row, col, row_t, col_t, sources = map(_i32, (row, col, row_t, col_t, sources))
    n = row.size - 1
    out = np.empty(n, np.float32)
    lib().orc_bc_ref_f32(n, _p(row), _p(col), _p(row_t), _p(col_t), _p(sources), sources.size, _p(out))
    return out


def block_scan_ref(values):
    v = _i32(values)
    out = np.empty_like(v)
    total = lib().orc_block_scan_ref(v.size, _p(v), _p(out))
    return out, total


def dp_model(groups, resident, lock=0, bar=1, skip=0, barriers=8, watchdog_ms=2000):
    ids = np.full(groups, -1, np.int32)
    p = C.c_int(0)
    bad = C.c_longlong(0)
    dead = lib().orc_dp_model(groups, resident, lock, bar, skip, barriers, watchdog_ms, C.byref(p),
                              _p(ids), C.byref(bad))
    return {"deadlock": bool(dead), "P": p.value, "ids": ids, "violations": bad.value}


MLP_PARAMS = 5508


def mlp_forward(w, x):
    """FP64 forward of the selector MLP for rows x[k][16]; w = parameter values."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 16)
    y = np.empty((x.shape[0], 4), dtype=np.float64)
    for r in range(x.shape[0]):
        xr = np.ascontiguousarray(x[r])
        yr = np.empty(4, dtype=np.float64)
        lib().orc_mlp_forward(_p(w), _p(xr), _p(yr))
        y[r] = yr
    return y


def mlp_grad(w, records):
    """(grad[5508], loss, used) of the trace loss in FP64."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    rec = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 20)
    g = np.empty(MLP_PARAMS, dtype=np.float64)
    loss = C.c_double(0)
    used = lib().orc_mlp_grad(_p(w), _p(rec), rec.shape[0], _p(g), C.byref(loss))
    return g, loss.value, used


def _mst(fn, row, col, w):
    row, col = _i32(row), _i32(col)
    w = np.ascontiguousarray(w, dtype=np.uint32)
    wt, ed, cp = C.c_ulonglong(0), C.c_int(0), C.c_int(0)
    fn(row.size - 1, _p(row), _p(col), _p(w), C.byref(wt), C.byref(ed), C.byref(cp))
    return wt.value, ed.value, cp.value


def mst_kruskal(row, col, w):
    """(forest weight, edges, components) of a symmetric CSR."""
    return _mst(lib().orc_mst_kruskal, row, col, w)


def mst_prim_tb(row, col, w):
    return _mst(lib().orc_mst_prim_tb, row, col, w)
