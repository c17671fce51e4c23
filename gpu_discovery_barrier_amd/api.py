"""ctypes binding of libgdb_amd.so (include/gdb_amd.h, include/gdb_graph.h).

The shared library is the product: every solver below calls its C ABI and
runs the HIP kernels.  Nothing here computes a result on the CPU; if the
library is missing, importing a solver raises instead of falling back.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libgdb_amd.so"

# Return codes (gdb_amd.h).
GDB_SUCCESS = 0
GDB_E_INVALID = -1
GDB_E_HIP = -2
GDB_E_TIMEOUT = -3
GDB_E_UNSAFE = -4
GDB_E_OVERFLOW = -5
GDB_E_NOMEM = -6

PROTOCOL_LOCKFREE, PROTOCOL_TICKET, PROTOCOL_SPIN = 0, 1, 2
BARRIER_XF, BARRIER_COUNTER, BARRIER_TREE, BARRIER_AUTO = 0, 1, 2, 3
SCHED_PERSISTENT, SCHED_RELAUNCH = 0, 1
VARIANT_PLAIN, VARIANT_TUNED = 0, 1

PANNOTIA_INF = 99999999
LONESTAR_INF = 1000000000


class GdbError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with code {code}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [
        ("wgs", C.c_int),
        ("groups", C.c_int),
        ("schedule", C.c_int),
        ("variant", C.c_int),
        ("barrier", C.c_int),
        ("relaunch_check", C.c_int),
        ("lds_queue", C.c_int),
    ]

    @classmethod
    def make(cls, wgs=256, groups=0, schedule=SCHED_PERSISTENT, variant=VARIANT_PLAIN,
             barrier=-1, relaunch_check=1, lds_queue=0) -> "Opts":
        """wgs = 0 takes the launch shape from the tuning table (tuned
        persistent variant, batched engine, MST)."""
        return cls(wgs, groups, schedule, variant, barrier, relaunch_check, lds_queue)


class TuneEntry(C.Structure):
    """gdb_tune_entry (include/gdb_amd.h): one row of the tuning table."""
    _fields_ = [
        ("solver", C.c_char * 16),
        ("n_max", C.c_longlong),
        ("wgs", C.c_int),
        ("groups", C.c_int),
        ("team_groups", C.c_int),
        ("delta", C.c_int),
        ("hops", C.c_int),
        ("lds_queue", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("participating", C.c_int),
        ("rounds", C.c_int),
        ("barriers", C.c_longlong),
        ("launches", C.c_longlong),
        ("kernel_ms", C.c_double),
        ("work_edges", C.c_longlong),
        ("work_items", C.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


TASK_BFS, TASK_SSSP, TASK_COLOR = 0, 1, 2


class Task(C.Structure):
    """gdb_task (include/gdb_amd.h): one (graph, algorithm) task of a batch."""
    _fields_ = [
        ("algo", C.c_int),
        ("n", C.c_int),
        ("m", C.c_int),
        ("row", C.c_void_p),
        ("col", C.c_void_p),
        ("w", C.c_void_p),
        ("source", C.c_int),
        ("node_value", C.c_void_p),
        ("out", C.c_void_p),
        ("delta", C.c_int),
        ("hops", C.c_int),
        ("owned", C.c_int),
        ("trow", C.c_void_p),
        ("tcol", C.c_void_p),
        ("tw", C.c_void_p),
        ("rounds", C.c_int),
        ("time_us", C.c_double),
        ("arm_iters", C.c_int * 4),
        ("work_items", C.c_longlong),
        ("work_edges", C.c_longlong),
    ]


# Schedule selector (gdb_selector): modes and arms.
SEL_NONE, SEL_FIXED, SEL_HEURISTIC, SEL_RANDOM, SEL_MLP = 0, 1, 2, 3, 4
ARM_PUSH, ARM_PULL, ARM_PUSH_NARROW, ARM_PULL_NARROW = 0, 1, 2, 3
MLP_IN, MLP_HIDDEN, MLP_OUT = 16, 64, 4
MLP_PARAMS = MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_OUT + MLP_OUT
TRACE_STRIDE = 20


class Selector(C.Structure):
    """gdb_selector (include/gdb_amd.h)."""
    _fields_ = [
        ("mode", C.c_int),
        ("fixed_arm", C.c_int),
        ("narrow_groups", C.c_int),
        ("seed", C.c_uint),
        ("weights", C.c_void_p),
        ("trace", C.c_void_p),
        ("trace_capacity", C.c_int),
        ("trace_count", C.c_void_p),
    ]


class HostCSRStruct(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("m", C.c_int),
        ("row", C.POINTER(C.c_int)),
        ("col", C.POINTER(C.c_int)),
        ("w", C.POINTER(C.c_int)),
    ]


_lib = None
P = C.c_void_p
I = C.c_int


def _declare(lib) -> None:
    sig = {
        "gdb_malloc": (I, [C.POINTER(P), C.c_size_t]),
        "gdb_free": (I, [P]),
        "gdb_memcpy_htod": (I, [P, P, C.c_size_t]),
        "gdb_memcpy_dtoh": (I, [P, P, C.c_size_t]),
        "gdb_memcpy_dtod": (I, [P, P, C.c_size_t]),
        "gdb_device_synchronize": (I, []),
        "gdb_device_count": (I, [C.POINTER(I)]),
        "gdb_ctx_create": (I, [C.POINTER(P)]),
        "gdb_ctx_destroy": (I, [P]),
        "gdb_ctx_configure": (I, [P, I, I]),
        "gdb_ctx_set_timeouts": (I, [P, C.c_double, C.c_double]),
        "gdb_ctx_set_xcd_only": (I, [P, I]),
        "gdb_ctx_init": (I, [P]),
        "gdb_ctx_init_skip": (I, [P, I]),
        "gdb_number_of_participating_groups": (I, [P]),
        "gdb_ctx_device_error": (I, [P]),
        "gdb_ctx_stream": (P, [P]),
        "gdb_ctx_device_ptr": (I, [P, C.POINTER(P)]),
        "gdb_launch_on_ctx": (I, [P, P, C.POINTER(P), I, I, C.c_size_t]),
        "gdb_time_protocol_kernel": (I, [P, P, C.POINTER(P), I, I, C.c_size_t, C.POINTER(C.c_double)]),
        "gdb_time_protocol": (I, [P, I, I, I, C.POINTER(C.c_double)]),
        "gdb_occupancy_test": (I, [P, I, I, I, I, C.POINTER(I)]),
        "gdb_max_resident_groups": (I, [I, I, C.POINTER(I)]),
        "gdb_device_query": (I, [C.c_char_p, C.c_size_t]),
        "gdb_barrier_bench": (I, [P, I, I, I, C.POINTER(C.c_double), C.POINTER(I)]),
        "gdb_barrier_check": (I, [P, I, I, I, C.POINTER(C.c_longlong), C.POINTER(I)]),
        "gdb_relaunch_bench": (I, [P, I, I, I, I, C.POINTER(C.c_double)]),
        "gdb_tune_load": (I, [C.c_char_p]),
        "gdb_tune_lookup": (I, [C.c_char_p, C.c_longlong, C.POINTER(TuneEntry)]),
        "gdb_probe_chain": (I, [P, I, C.c_size_t, I, I, C.POINTER(C.c_double)]),
        "gdb_probe_atomics": (I, [P, I, I, I, I, C.POINTER(C.c_double)]),
        "gdb_probe_pingpong": (I, [P, I, I, I, I, C.POINTER(C.c_double), C.POINTER(I)]),
        "gdb_color": (I, [P, I, I, P, P, P, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_csr_free": (None, [C.POINTER(HostCSRStruct)]),
        "gdb_gen_rmat": (I, [I, I, C.c_double, C.c_double, C.c_double, C.c_uint64, I, I,
                             C.POINTER(HostCSRStruct)]),
        "gdb_gen_grid": (I, [I, I, C.c_double, C.c_uint64, I, I, C.POINTER(HostCSRStruct)]),
        "gdb_gen_gnp": (I, [I, C.c_double, C.c_uint64, I, I, C.POINTER(HostCSRStruct)]),
        "gdb_gen_permutation": (I, [I, C.c_uint64, P]),
        "gdb_gen_uniform_ints": (I, [C.c_longlong, I, I, C.c_uint64, P]),
        "gdb_csr_transpose": (I, [C.POINTER(HostCSRStruct), C.POINTER(HostCSRStruct)]),
    }
    optional = {
        "gdb_csr_relabel": (I, [C.POINTER(HostCSRStruct), P, C.POINTER(HostCSRStruct)]),
        "gdb_order_by_degree": (I, [C.POINTER(HostCSRStruct), P]),
        "gdb_mis": (I, [P, I, I, P, P, P, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_color_maxmin": (I, [P, I, I, P, P, P, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_sssp_pull": (I, [P, I, I, P, P, P, I, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_bc": (I, [P, I, I, P, P, P, P, P, I, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_bfs_worklist": (I, [P, I, I, P, P, I, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_sssp_worklist": (I, [P, I, I, P, P, P, I, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_sssp_worklist_noalloc": (I, [P, I, I, P, P, P, I, P, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_verify_distances": (I, [P, I, I, P, P, P, P, C.POINTER(C.c_longlong)]),
        "gdb_mst": (I, [P, I, I, P, P, P, C.POINTER(C.c_ulonglong), C.POINTER(I), C.POINTER(I),
                        C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_batch_solve": (I, [P, C.POINTER(Task), I, I, C.POINTER(Opts), C.POINTER(Stats)]),
        "gdb_batch_solve_sel": (I, [P, C.POINTER(Task), I, I, C.POINTER(Opts), C.POINTER(Selector),
                                    C.POINTER(Stats)]),
        "gdb_batch_graph_create": (I, [P, C.POINTER(Task), I, I, C.POINTER(Opts), C.POINTER(Selector),
                                       C.POINTER(P)]),
        "gdb_batch_graph_launch": (I, [P, C.POINTER(Stats)]),
        "gdb_batch_graph_destroy": (I, [P]),
        "gdb_mlp_forward": (I, [P, P, P, I, P]),
        "gdb_mlp_grad": (I, [P, P, P, I, P, C.POINTER(C.c_float), C.POINTER(I)]),
        "gdb_mlp_to_bf16": (I, [P, P, P, I]),
        "gdb_mlp_train": (I, [P, P, P, P, P, P, I, I, I, C.c_float, C.c_uint, C.POINTER(C.c_float)]),
        "gdb_load_metis": (I, [C.c_char_p, I, C.POINTER(HostCSRStruct)]),
        "gdb_load_dimacs": (I, [C.c_char_p, I, I, C.POINTER(HostCSRStruct)]),
        "gdb_load_galois_gr": (I, [C.c_char_p, C.POINTER(HostCSRStruct)]),
        "gdb_write_galois_gr": (I, [C.c_char_p, C.POINTER(HostCSRStruct)]),
        "gdb_load_edges": (I, [C.c_char_p, C.POINTER(HostCSRStruct)]),
        "gdb_csr_validate": (I, [C.POINTER(HostCSRStruct)]),
    }
    for name, (res, args) in {**sig, **optional}.items():
        if name in optional and not hasattr(lib, name):
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Loads libgdb_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m gpu_discovery_barrier_amd.build`")
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def check(rc: int, what: str) -> None:
    if rc != GDB_SUCCESS:
        raise GdbError(rc, what)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().gdb_device_count(C.byref(n)), "gdb_device_count")
    return n.value


# ---------------------------------------------------------------------------
# Device memory
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A device allocation made by gdb_malloc."""

    def __init__(self, nbytes: int, dtype=np.int32, count: int | None = None):
        self.ptr = P()
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.count = count if count is not None else self.nbytes // self.dtype.itemsize
        check(lib().gdb_malloc(C.byref(self.ptr), max(self.nbytes, 4)), "gdb_malloc")

    @classmethod
    def empty(cls, count: int, dtype=np.int32) -> "DeviceBuffer":
        return cls(int(count) * np.dtype(dtype).itemsize, dtype, int(count))

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes, arr.dtype, arr.size)
        if arr.nbytes:
            check(lib().gdb_memcpy_htod(buf.ptr, arr.ctypes.data_as(P), arr.nbytes), "gdb_memcpy_htod")
        return buf

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        if out.nbytes:
            check(lib().gdb_memcpy_dtoh(out.ctypes.data_as(P), self.ptr, out.nbytes), "gdb_memcpy_dtoh")
        return out

    def free(self) -> None:
        if self.ptr:
            lib().gdb_free(self.ptr)
            self.ptr = P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Host graphs (generators run in the library's host code)
# ---------------------------------------------------------------------------
@dataclass
class CSR:
    n: int
    row: np.ndarray  # int32[n+1]
    col: np.ndarray  # int32[m]
    w: np.ndarray    # int32[m]

    @property
    def m(self) -> int:
        return int(self.col.size)

    def transpose(self) -> "CSR":
        src = _to_struct(self)
        out = HostCSRStruct()
        check(lib().gdb_csr_transpose(C.byref(src), C.byref(out)), "gdb_csr_transpose")
        return _from_struct(out)

    def degrees(self) -> np.ndarray:
        return np.diff(self.row)

    def relabel(self, new_id: np.ndarray) -> "CSR":
        """The graph with vertex v renamed new_id[v] (gdb_csr_relabel)."""
        src = _to_struct(self)
        ids = np.ascontiguousarray(new_id, np.int32)
        out = HostCSRStruct()
        check(lib().gdb_csr_relabel(C.byref(src), ids.ctypes.data_as(P), C.byref(out)), "gdb_csr_relabel")
        return _from_struct(out)

    def order_by_degree(self) -> np.ndarray:
        """new_id putting vertices in decreasing degree order (ties by id)."""
        src = _to_struct(self)
        ids = np.empty(self.n, np.int32)
        check(lib().gdb_order_by_degree(C.byref(src), ids.ctypes.data_as(P)), "gdb_order_by_degree")
        return ids


def _from_struct(s: HostCSRStruct) -> CSR:
    n, m = s.n, s.m
    row = np.ctypeslib.as_array(s.row, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(s.col, shape=(max(m, 1),))[:m].copy()
    w = np.ctypeslib.as_array(s.w, shape=(max(m, 1),))[:m].copy()
    lib().gdb_csr_free(C.byref(s))
    return CSR(n, row, col, w)


def _to_struct(g: CSR) -> HostCSRStruct:
    s = HostCSRStruct()
    s.n, s.m = g.n, g.m
    g._keep = [np.ascontiguousarray(g.row, np.int32), np.ascontiguousarray(g.col, np.int32),
               np.ascontiguousarray(g.w, np.int32)]
    s.row = g._keep[0].ctypes.data_as(C.POINTER(C.c_int))
    s.col = g._keep[1].ctypes.data_as(C.POINTER(C.c_int))
    s.w = g._keep[2].ctypes.data_as(C.POINTER(C.c_int))
    return s


def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1,
             wlo: int = 1, whi: int = 1) -> CSR:
    out = HostCSRStruct()
    check(lib().gdb_gen_rmat(scale, edge_factor, a, b, c, seed, wlo, whi, C.byref(out)), "gdb_gen_rmat")
    return _from_struct(out)


def gen_grid(rows: int, cols: int, remove_frac: float = 0.0, seed: int = 2, wlo: int = 1,
             whi: int = 1) -> CSR:
    out = HostCSRStruct()
    check(lib().gdb_gen_grid(rows, cols, remove_frac, seed, wlo, whi, C.byref(out)), "gdb_gen_grid")
    return _from_struct(out)


def gen_gnp(n: int, avg_degree: float = 8.0, seed: int = 4, wlo: int = 1, whi: int = 1) -> CSR:
    out = HostCSRStruct()
    check(lib().gdb_gen_gnp(n, avg_degree, seed, wlo, whi, C.byref(out)), "gdb_gen_gnp")
    return _from_struct(out)


def permutation(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    check(lib().gdb_gen_permutation(n, seed, out.ctypes.data_as(P)), "gdb_gen_permutation")
    return out


def uniform_ints(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    out = np.empty(count, np.int32)
    check(lib().gdb_gen_uniform_ints(count, lo, hi, seed, out.ctypes.data_as(P)), "gdb_gen_uniform_ints")
    return out


def pannotia_values(n: int) -> np.ndarray:
    """node_value[i] = i / (float)(n + 1) (color_gb.cpp:152, mis_gb.cpp:151)."""
    return (np.arange(n, dtype=np.float32) / np.float32(n + 1)).astype(np.float32)


# ---------------------------------------------------------------------------
# Device graphs and the discovery context
# ---------------------------------------------------------------------------
def validate_csr(g: CSR) -> None:
    """Rejects a CSR whose offsets or destinations would index out of range."""
    row, col = np.asarray(g.row), np.asarray(g.col)
    if row.size != g.n + 1 or row[0] != 0 or row[-1] != col.size or (np.diff(row) < 0).any():
        raise ValueError("malformed CSR offsets")
    if col.size and (col.min() < 0 or col.max() >= g.n):
        raise ValueError("CSR destination out of range")


class DeviceCSR:
    def __init__(self, g: CSR, weights: bool = True):
        validate_csr(g)
        self.n, self.m = g.n, g.m
        self.row = DeviceBuffer.from_numpy(g.row.astype(np.int32))
        self.col = DeviceBuffer.from_numpy(g.col.astype(np.int32))
        self.w = DeviceBuffer.from_numpy(g.w.astype(np.int32)) if weights else None


class Context:
    """A discovery context (gdb_ctx): device protocol/barrier state + stream."""

    def __init__(self, lock_kind: int = PROTOCOL_LOCKFREE, bar_kind: int = BARRIER_AUTO):
        self.h = P()
        check(lib().gdb_ctx_create(C.byref(self.h)), "gdb_ctx_create")
        self.configure(lock_kind, bar_kind)

    def configure(self, lock_kind: int, bar_kind: int) -> None:
        check(lib().gdb_ctx_configure(self.h, lock_kind, bar_kind), "gdb_ctx_configure")
        check(lib().gdb_ctx_init(self.h), "gdb_ctx_init")

    def set_xcd_only(self, xcd: int) -> None:
        check(lib().gdb_ctx_set_xcd_only(self.h, xcd), "gdb_ctx_set_xcd_only")
        check(lib().gdb_ctx_init(self.h), "gdb_ctx_init")

    def set_timeouts(self, wait_us: float, quiet_us: float) -> None:
        check(lib().gdb_ctx_set_timeouts(self.h, wait_us, quiet_us), "gdb_ctx_set_timeouts")

    def init(self, skip: bool = False) -> None:
        check(lib().gdb_ctx_init_skip(self.h, int(skip)), "gdb_ctx_init_skip")

    def participating(self) -> int:
        rc = lib().gdb_number_of_participating_groups(self.h)
        if rc < 0:
            raise GdbError(rc, "gdb_number_of_participating_groups")
        return rc

    def close(self) -> None:
        if self.h:
            lib().gdb_ctx_destroy(self.h)
            self.h = P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- protocol / barrier tools ---------------------------------------
    def time_protocol(self, upper_bound: int, wgs: int, lms: int = 0) -> float:
        ns = C.c_double(0)
        check(lib().gdb_time_protocol(self.h, upper_bound, wgs, lms, C.byref(ns)), "gdb_time_protocol")
        return ns.value

    def occupancy_test(self, wgc: int, wgs: int, lms: int = 0, prot: bool = True) -> int:
        p = C.c_int(0)
        check(lib().gdb_occupancy_test(self.h, wgc, wgs, lms, int(prot), C.byref(p)), "gdb_occupancy_test")
        return p.value

    def barrier_bench(self, wgs: int, groups: int, iters: int) -> tuple[float, int]:
        ms, p = C.c_double(0), C.c_int(0)
        check(lib().gdb_barrier_bench(self.h, wgs, groups, iters, C.byref(ms), C.byref(p)),
              "gdb_barrier_bench")
        return ms.value, p.value

    def barrier_check(self, wgs: int, groups: int, iters: int) -> tuple[int, int]:
        bad, p = C.c_longlong(0), C.c_int(0)
        check(lib().gdb_barrier_check(self.h, wgs, groups, iters, C.byref(bad), C.byref(p)),
              "gdb_barrier_check")
        return bad.value, p.value

    def probe_chain(self, kind: int, nbytes: int, fine: bool, steps: int) -> float:
        """ns per dependent access (kind 0 load, 1 atomic add, 2 agent-scope load)."""
        ns = C.c_double(0)
        check(lib().gdb_probe_chain(self.h, kind, nbytes, int(fine), steps, C.byref(ns)), "gdb_probe_chain")
        return ns.value

    def probe_atomics(self, bits: int, groups: int, wgs: int, iters: int) -> float:
        """ns per step of dependent returning atomic-min operations (every lane)."""
        ns = C.c_double(0)
        check(lib().gdb_probe_atomics(self.h, bits, groups, wgs, iters, C.byref(ns)), "gdb_probe_atomics")
        return ns.value

    def probe_pingpong(self, mode: int, fine: bool, same_xcd: bool, iters: int) -> tuple[float, list]:
        """(ns per handoff between two groups, their XCD ids)."""
        ns, x = C.c_double(0), (C.c_int * 2)()
        check(lib().gdb_probe_pingpong(self.h, mode, int(fine), int(same_xcd), iters, C.byref(ns), x),
              "gdb_probe_pingpong")
        return ns.value, [x[0], x[1]]

    def relaunch_bench(self, wgs: int, groups: int, iters: int, mode: int = 0) -> float:
        ms = C.c_double(0)
        check(lib().gdb_relaunch_bench(self.h, wgs, groups, iters, mode, C.byref(ms)), "gdb_relaunch_bench")
        return ms.value


def tune_load(path: str | None = None) -> int:
    """Replaces the library's tuning table with a file's rows (None: the
    default file); returns the number of rows."""
    rc = lib().gdb_tune_load(None if path is None else str(path).encode())
    if rc < 0:
        raise GdbError(rc, "gdb_tune_load")
    return rc


def tune_lookup(solver: str, n: int) -> dict | None:
    """The tuning-table row a solver of graph size n uses, or None."""
    e = TuneEntry()
    if lib().gdb_tune_lookup(solver.encode(), int(n), C.byref(e)) != GDB_SUCCESS:
        return None
    return {k: (getattr(e, k).decode() if k == "solver" else getattr(e, k)) for k, _ in TuneEntry._fields_}


def max_resident_groups(wgs: int, lms: int = 0) -> int:
    out = C.c_int(0)
    check(lib().gdb_max_resident_groups(wgs, lms, C.byref(out)), "gdb_max_resident_groups")
    return out.value


def device_query() -> str:
    buf = C.create_string_buffer(4096)
    check(lib().gdb_device_query(buf, 4096), "gdb_device_query")
    return buf.value.decode()


# ---------------------------------------------------------------------------
# Solvers
# ---------------------------------------------------------------------------
def _opts(opts: Opts | dict | None) -> Opts:
    if opts is None:
        return Opts.make()
    if isinstance(opts, dict):
        return Opts.make(**opts)
    return opts


def color(ctx: Context, g: DeviceCSR, node_value: DeviceBuffer, opts=None, out: DeviceBuffer | None = None):
    out = out or DeviceBuffer.empty(g.n, np.int32)
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_color(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr, node_value.ptr, out.ptr, C.byref(o),
                          C.byref(st)), "gdb_color")
    return out.to_numpy(), st


def color_maxmin(ctx: Context, g: DeviceCSR, node_value: DeviceBuffer, opts=None,
                 out: DeviceBuffer | None = None):
    """Max-min colouring (colours 2r-1 / 2r for round r's local maxima / minima)."""
    out = out or DeviceBuffer.empty(g.n, np.int32)
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_color_maxmin(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr, node_value.ptr, out.ptr, C.byref(o),
                                 C.byref(st)), "gdb_color_maxmin")
    return out.to_numpy(), st


def mis(ctx: Context, g: DeviceCSR, node_value: DeviceBuffer, opts=None, out: DeviceBuffer | None = None):
    out = out or DeviceBuffer.empty(g.n, np.int32)
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_mis(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr, node_value.ptr, out.ptr, C.byref(o),
                        C.byref(st)), "gdb_mis")
    return out.to_numpy(), st


def sssp_pull(ctx: Context, gt: DeviceCSR, source: int = 0, opts=None, out: DeviceBuffer | None = None):
    out = out or DeviceBuffer.empty(gt.n, np.int32)
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_sssp_pull(ctx.h, gt.n, gt.m, gt.row.ptr, gt.col.ptr, gt.w.ptr, source, out.ptr,
                              C.byref(o), C.byref(st)), "gdb_sssp_pull")
    return out.to_numpy(), st


def bc(ctx: Context, g: DeviceCSR, gt: DeviceCSR, sources: np.ndarray, opts=None,
       out: DeviceBuffer | None = None):
    out = out or DeviceBuffer.empty(g.n, np.float32)
    src = DeviceBuffer.from_numpy(np.asarray(sources, np.int32))
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_bc(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr, gt.row.ptr, gt.col.ptr, src.ptr,
                       int(len(sources)), out.ptr, C.byref(o), C.byref(st)), "gdb_bc")
    return out.to_numpy(), st


def bfs_worklist(ctx: Context, g: DeviceCSR, source: int = 0, opts=None, out: DeviceBuffer | None = None):
    out = out or DeviceBuffer.empty(g.n, np.uint32)
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_bfs_worklist(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr, source, out.ptr, C.byref(o),
                                 C.byref(st)), "gdb_bfs_worklist")
    return out.to_numpy(), st


def sssp_worklist(ctx: Context, g: DeviceCSR, source: int = 0, opts=None,
                  out: DeviceBuffer | None = None):
    out = out or DeviceBuffer.empty(g.n, np.uint32)
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_sssp_worklist(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr, g.w.ptr, source, out.ptr,
                                  C.byref(o), C.byref(st)), "gdb_sssp_worklist")
    return out.to_numpy(), st


def mst(ctx: Context, g: DeviceCSR, opts=None):
    """Minimum spanning forest of a symmetric graph: (weight, edges, components, Stats)."""
    if g.w is None:
        raise ValueError("mst needs edge weights")
    wt, ed, cp = C.c_ulonglong(0), C.c_int(0), C.c_int(0)
    st = Stats()
    o = _opts(opts)
    check(lib().gdb_mst(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr, g.w.ptr, C.byref(wt), C.byref(ed), C.byref(cp),
                        C.byref(o), C.byref(st)), "gdb_mst")
    return wt.value, ed.value, cp.value, st


def verify_distances(ctx: Context, g: DeviceCSR, dist: DeviceBuffer, weighted: bool) -> int:
    nerr = C.c_longlong(0)
    check(lib().gdb_verify_distances(ctx.h, g.n, g.m, g.row.ptr, g.col.ptr,
                                     g.w.ptr if weighted else None, dist.ptr, C.byref(nerr)),
          "gdb_verify_distances")
    return nerr.value


# ---------------------------------------------------------------------------
# Batched engine
# ---------------------------------------------------------------------------
def make_task(algo: int, g: "DeviceCSR", out: DeviceBuffer, source: int = 0,
              node_value: DeviceBuffer | None = None, delta: int = 0, hops: int = 0,
              transpose: "DeviceCSR | None" = None, owned: bool = False) -> Task:
    """`transpose` is the in-edge CSR (the graph itself when it is symmetric);
    it enables the pull arms of the schedule selector.  `owned` asks for the
    owner-partitioned solve (each group of the team keeps a contiguous vertex
    range in LDS; see gdb_task.owned).  The task holds references to the
    device objects it points into, so they stay allocated while the task is
    alive."""
    t = Task()
    t._keep = (g, out, node_value, transpose)
    t.algo, t.n, t.m = algo, g.n, g.m
    t.row, t.col = g.row.ptr, g.col.ptr
    t.w = g.w.ptr if g.w is not None else None
    t.source = source
    t.node_value = node_value.ptr if node_value is not None else None
    t.out = out.ptr
    t.delta = delta
    t.hops = hops
    t.owned = 1 if owned else 0
    if transpose is not None:
        if transpose.n != g.n or transpose.m != g.m:
            raise ValueError("transpose must have the graph's n and m")
        t.trow, t.tcol = transpose.row.ptr, transpose.col.ptr
        t.tw = transpose.w.ptr if transpose.w is not None else None
    t.rounds = 0
    return t


def make_selector(mode: int = SEL_NONE, fixed_arm: int = 0, narrow_groups: int = 1, seed: int = 0,
                  weights: "DeviceBuffer | None" = None, trace: "DeviceBuffer | None" = None,
                  trace_count: "DeviceBuffer | None" = None) -> Selector:
    s = Selector()
    # The structure holds raw device pointers; keep their owners alive with it.
    s._keep = (weights, trace, trace_count)
    s.mode, s.fixed_arm, s.narrow_groups, s.seed = mode, fixed_arm, narrow_groups, seed & 0xFFFFFFFF
    s.weights = weights.ptr if weights is not None else None
    if trace is not None:
        if trace_count is None:
            raise ValueError("a trace buffer needs a trace_count buffer")
        s.trace = trace.ptr
        s.trace_capacity = trace.count // TRACE_STRIDE
        s.trace_count = trace_count.ptr
    return s


def batch_solve(ctx: Context, tasks: list, team_groups: int = 1, opts=None, selector: Selector | None = None):
    """Runs the tasks in one persistent launch; returns (rounds per task, Stats).
    With a selector, BFS / SSSP iterations pick their schedule arm per
    iteration; each task's arm_iters counts the iterations per arm."""
    arr = (Task * len(tasks))(*tasks)
    st = Stats()
    o = _opts(opts)
    if selector is None:
        check(lib().gdb_batch_solve(ctx.h, arr, len(tasks), team_groups, C.byref(o), C.byref(st)),
              "gdb_batch_solve")
    else:
        check(lib().gdb_batch_solve_sel(ctx.h, arr, len(tasks), team_groups, C.byref(o), C.byref(selector),
                                        C.byref(st)), "gdb_batch_solve_sel")
    for i, t in enumerate(tasks):
        t.rounds, t.time_us = arr[i].rounds, arr[i].time_us
        t.arm_iters[:] = arr[i].arm_iters[:]
        t.work_items, t.work_edges = arr[i].work_items, arr[i].work_edges
    return [arr[i].rounds for i in range(len(tasks))], st


class BatchGraph:
    """A batch recorded once as a HIP graph and replayed (gdb_batch_graph_*).
    Keeps its own copy of the task array; replays update those tasks'
    rounds / time_us / arm_iters and the caller's Task objects."""

    def __init__(self, ctx: Context, tasks: list, team_groups: int = 1, opts=None,
                 selector: "Selector | None" = None):
        self.tasks = tasks
        # The recorded kernel arguments point into the tasks' and the
        # selector's device buffers: both stay referenced by the graph.
        self.selector = selector
        self.arr = (Task * len(tasks))(*tasks)
        self.h = P()
        o = _opts(opts)
        check(lib().gdb_batch_graph_create(ctx.h, self.arr, len(tasks), team_groups, C.byref(o),
                                           C.byref(selector) if selector is not None else None, C.byref(self.h)),
              "gdb_batch_graph_create")

    def launch(self):
        st = Stats()
        check(lib().gdb_batch_graph_launch(self.h, C.byref(st)), "gdb_batch_graph_launch")
        for i, t in enumerate(self.tasks):
            t.rounds, t.time_us = self.arr[i].rounds, self.arr[i].time_us
            t.arm_iters[:] = self.arr[i].arm_iters[:]
            t.work_items, t.work_edges = self.arr[i].work_items, self.arr[i].work_edges
        return [self.arr[i].rounds for i in range(len(self.tasks))], st

    def close(self) -> None:
        if self.h:
            lib().gdb_batch_graph_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Selector MLP (forward / backward / Adam on the device)
# ---------------------------------------------------------------------------
def bf16_round(x: np.ndarray) -> np.ndarray:
    """fp32 -> bf16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    nan = np.isnan(np.asarray(x, dtype=np.float32))
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def mlp_init(seed: int = 5) -> np.ndarray:
    """He-uniform initial parameters (fp32) in the library's layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((MLP_IN, MLP_HIDDEN), (MLP_HIDDEN, MLP_HIDDEN), (MLP_HIDDEN, MLP_OUT)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, size=fan_in * fan_out))
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == MLP_PARAMS
    return p


def mlp_forward(ctx: Context, w_bf16: DeviceBuffer, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
    dx = DeviceBuffer.from_numpy(x)
    dy = DeviceBuffer.empty(x.shape[0] * MLP_OUT, np.float32)
    check(lib().gdb_mlp_forward(ctx.h, w_bf16.ptr, dx.ptr, x.shape[0], dy.ptr), "gdb_mlp_forward")
    return dy.to_numpy().reshape(-1, MLP_OUT)


def mlp_grad(ctx: Context, w_bf16: DeviceBuffer, records: DeviceBuffer, count: int):
    """(gradient[MLP_PARAMS], loss, records used) of the trace loss."""
    g = DeviceBuffer.empty(MLP_PARAMS, np.float32)
    loss, used = C.c_float(0), C.c_int(0)
    check(lib().gdb_mlp_grad(ctx.h, w_bf16.ptr, records.ptr, count, g.ptr, C.byref(loss), C.byref(used)),
          "gdb_mlp_grad")
    return g.to_numpy(), loss.value, used.value


def mlp_train(ctx: Context, params: np.ndarray, records: DeviceBuffer, count: int, steps: int = 2000,
              batch: int = 256, lr: float = 2e-3, seed: int = 5):
    """Adam on the device; returns (fp32 params, bf16 weights as DeviceBuffer, final loss)."""
    dp = DeviceBuffer.from_numpy(np.asarray(params, dtype=np.float32))
    dm = DeviceBuffer.from_numpy(np.zeros(MLP_PARAMS, np.float32))
    dv = DeviceBuffer.from_numpy(np.zeros(MLP_PARAMS, np.float32))
    dw = DeviceBuffer.empty(MLP_PARAMS, np.uint16)
    loss = C.c_float(0)
    check(lib().gdb_mlp_train(ctx.h, dp.ptr, dm.ptr, dv.ptr, dw.ptr, records.ptr, count, steps, batch,
                              C.c_float(lr), seed & 0xFFFFFFFF, C.byref(loss)), "gdb_mlp_train")
    return dp.to_numpy(), dw, loss.value


# ---------------------------------------------------------------------------
# Graph files
# ---------------------------------------------------------------------------
def _load(fn, what, *args) -> CSR:
    out = HostCSRStruct()
    check(fn(*args, C.byref(out)), what)
    return _from_struct(out)


def load_metis(path: str, directed: bool = False) -> CSR:
    return _load(lib().gdb_load_metis, "gdb_load_metis", str(path).encode(), int(directed))


def load_dimacs(path: str, directed: bool = True, transpose: bool = False) -> CSR:
    return _load(lib().gdb_load_dimacs, "gdb_load_dimacs", str(path).encode(), int(directed), int(transpose))


def load_galois_gr(path: str) -> CSR:
    return _load(lib().gdb_load_galois_gr, "gdb_load_galois_gr", str(path).encode())


def load_edges(path: str) -> CSR:
    return _load(lib().gdb_load_edges, "gdb_load_edges", str(path).encode())


def write_galois_gr(path: str, g: CSR) -> None:
    s = _to_struct(g)
    check(lib().gdb_write_galois_gr(str(path).encode(), C.byref(s)), "gdb_write_galois_gr")
