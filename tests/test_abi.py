"""The C-ABI library loads and exports every symbol that include/*.h declares
(no compute calls: this runs without a GPU)."""
import re
import subprocess
from pathlib import Path

import gpu_discovery_barrier_amd as gdb

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void\s*\*?|long long)\s+(gdb_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) > 20
    lib = gdb.lib()
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", str(gdb.api.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert declared_symbols() <= exported


def test_generators_are_deterministic_and_symmetric():
    a = gdb.gen_rmat(12, 16, seed=7, wlo=1, whi=100)
    b = gdb.gen_rmat(12, 16, seed=7, wlo=1, whi=100)
    assert (a.row == b.row).all() and (a.col == b.col).all() and (a.w == b.w).all()
    t = a.transpose()
    assert (t.row == a.row).all() and (t.col == a.col).all() and (t.w == a.w).all()
    src = a.col.copy()
    rows = __import__("numpy").repeat(__import__("numpy").arange(a.n), a.degrees())
    assert not (rows == src).any()  # no self-loops
    for i in range(0, a.n, 97):  # sorted, no duplicates
        seg = a.col[a.row[i]:a.row[i + 1]]
        assert (seg[1:] > seg[:-1]).all()
    assert 1 <= a.w.min() and a.w.max() <= 100


def test_grid_removes_exactly_five_percent():
    g = gdb.gen_grid(64, 48, 0.05, seed=2)
    undirected = 64 * 47 + 63 * 48
    assert g.m == 2 * (undirected - round(0.05 * undirected))


def test_gnp_average_degree():
    g = gdb.gen_gnp(20000, 8.0, seed=4)
    assert abs(g.m / g.n - 8.0) < 0.3


def test_task_keeps_its_device_objects_alive():
    # make_task copies raw device pointers into the C struct; the Python
    # objects that own that memory must live as long as the task, or a
    # launch would read freed memory.  Stand-ins with the same attributes
    # (no device allocation here).
    import gc
    import weakref

    class Buf:
        def __init__(self, ptr):
            self.ptr = ptr

    class Graph:
        def __init__(self):
            self.n, self.m = 4, 6
            self.row, self.col, self.w = Buf(16), Buf(32), Buf(48)

    g, out, nv = Graph(), Buf(64), Buf(80)
    refs = [weakref.ref(x) for x in (g, out, nv)]
    t = gdb.make_task(gdb.TASK_COLOR, g, out, node_value=nv)
    assert (t.row, t.col, t.out, t.node_value) == (16, 32, 64, 80)
    del g, out, nv
    gc.collect()
    assert all(r() is not None for r in refs)
    del t
    gc.collect()
    assert all(r() is None for r in refs)


def test_tuning_table_load_lookup_and_rejects_malformed(tmp_path):
    # The table the autotuner writes and every solver consults when called
    # with wgs = 0: the row for the solver with the smallest n_max >= n.
    from gpu_discovery_barrier_amd import autotune

    t = tmp_path / "tune.txt"
    autotune.write_table([autotune.row("batch", 1000, 128, 8, 4, 50, 16, 512),
                          autotune.row("batch", 1 << 30, 256, 64, 32, 4096, 256, 2048),
                          autotune.row("color", 5000, 64, 16)], t, header="test")
    try:
        assert gdb.tune_load(str(t)) == 3
        assert gdb.tune_lookup("batch", 999)["groups"] == 8
        assert gdb.tune_lookup("batch", 1000)["lds_queue"] == 512
        assert gdb.tune_lookup("batch", 1001)["wgs"] == 256
        assert gdb.tune_lookup("color", 6000) is None
        assert gdb.tune_lookup("mis", 10) is None
        assert autotune.lookup("batch", 1001, t)["team_groups"] == 32  # the Python reader agrees
        bad = tmp_path / "bad.txt"
        bad.write_text("batch 100 100 8 1 0 0 0\n")  # wgs 100 is not a supported size
        import pytest

        with pytest.raises(gdb.GdbError):
            gdb.tune_load(str(bad))
        assert gdb.tune_lookup("batch", 10) is None  # a rejected file leaves the table empty
    finally:
        gdb.tune_load(None)  # back to the default table beside the library
    assert gdb.tune_lookup("batch", 2048 * 2048) == autotune.lookup("batch", 2048 * 2048)


def test_committed_tuning_table_is_well_formed():
    from gpu_discovery_barrier_amd import autotune

    rows = autotune.read_table()
    assert {r["solver"] for r in rows} >= {"batch", "color", "mis", "bc"}
    for r in rows:
        assert r["wgs"] in (64, 128, 256, 512, 1024)
        assert r["lds_queue"] == 0 or 256 <= r["lds_queue"] <= 4096


def test_relabel_by_degree_is_an_isomorphism():
    import numpy as np

    g = gdb.gen_rmat(10, 16, seed=3, wlo=1, whi=50)
    nid = g.order_by_degree()
    assert sorted(nid.tolist()) == list(range(g.n))
    deg = g.degrees()
    assert (np.diff(deg[np.argsort(nid)]) <= 0).all()  # new ids in decreasing degree
    r = g.relabel(nid)
    assert r.m == g.m
    # Every edge (u, v, w) appears as (new_id[u], new_id[v], w), rows sorted.
    src = np.repeat(np.arange(g.n), deg)
    want = sorted(zip(nid[src].tolist(), nid[g.col].tolist(), g.w.tolist()))
    rsrc = np.repeat(np.arange(r.n), r.degrees())
    assert sorted(zip(rsrc.tolist(), r.col.tolist(), r.w.tolist())) == want
    for x in range(0, r.n, 37):
        seg = r.col[r.row[x]:r.row[x + 1]]
        assert (seg[1:] >= seg[:-1]).all()
    import pytest

    with pytest.raises(gdb.GdbError):
        g.relabel(np.zeros(g.n, np.int32))  # not a permutation
