"""Graph file loaders (include/gdb_graph.h) against hand-derived CSR arrays,
against the reference's own Pannotia parsers compiled from its sources into
oracle/_ref (when the reference tree is present), and by write/read round
trips."""
import ctypes as C
import os
from pathlib import Path

import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb

ROOT = Path(__file__).resolve().parents[1]
REF_LIB = ROOT / "oracle" / "_ref" / "libpanparse.so"
REF_SRC = Path("/root/reference/code/experiments/pannotia/graph_parser/parse.cpp")


def ref_parser():
    if not REF_LIB.exists():
        if not REF_SRC.exists():
            return None
        import subprocess
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    lib = C.CDLL(str(REF_LIB))
    IP = C.POINTER(C.c_int)
    lib.ref_parse.argtypes = [C.c_char_p, C.c_int, C.c_int, IP, IP, C.POINTER(IP), C.POINTER(IP), C.POINTER(IP)]
    lib.ref_parse.restype = C.c_int
    return lib


def ref_load(lib, path, kind, directed):
    n, m = C.c_int(), C.c_int()
    row, col, w = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    assert lib.ref_parse(str(path).encode(), kind, directed, C.byref(n), C.byref(m), C.byref(row),
                         C.byref(col), C.byref(w)) == 0
    N, M = n.value, m.value
    return (np.ctypeslib.as_array(row, (N + 1,)).copy(), np.ctypeslib.as_array(col, (M,)).copy(),
            np.ctypeslib.as_array(w, (M,)).copy())


METIS = """% a comment
4 4
2 3
1 3 4
1 2
2
"""
# Vertex 1: {2,3}; 2: {1,3,4}; 3: {1,2}; 4: {2}  (1-based) -> 0-based rows.
METIS_ROW = [0, 2, 5, 7, 8]
METIS_COL = [1, 2, 0, 2, 3, 0, 1, 1]

DIMACS = """c tiny
p sp 4 4
a 1 2 7
a 2 3 1
a 1 3 5
a 4 1 2
"""


def test_metis_hand(tmp_path):
    f = tmp_path / "g.graph"
    f.write_text(METIS)
    g = gdb.load_metis(f, directed=False)
    assert g.n == 4 and g.row.tolist() == METIS_ROW and g.col.tolist() == METIS_COL


def test_metis_empty_row_is_correct_csr(tmp_path):
    # Vertex 2 has no neighbours: the reference shifts later rows; the loader
    # keeps a correct CSR.
    f = tmp_path / "g.graph"
    f.write_text("3 1\n3\n\n1\n")
    g = gdb.load_metis(f, directed=False)
    assert g.row.tolist() == [0, 1, 1, 2] and g.col.tolist() == [2, 0]


def test_dimacs_hand(tmp_path):
    f = tmp_path / "g.gr"
    f.write_text(DIMACS)
    d = gdb.load_dimacs(f, directed=True)
    # stable sort by source: row 0 = arcs 1->2 (7), 1->3 (5); row 1 = 2->3; row 3 = 4->1
    assert d.row.tolist() == [0, 2, 3, 3, 4]
    assert d.col.tolist() == [1, 2, 2, 0] and d.w.tolist() == [7, 5, 1, 2]
    u = gdb.load_dimacs(f, directed=False)
    # every arc also reversed, in file order, then stable-sorted by row
    assert u.row.tolist() == [0, 3, 5, 7, 8]
    assert u.col.tolist() == [1, 2, 3, 0, 2, 1, 0, 0]
    t = gdb.load_dimacs(f, directed=True, transpose=True)
    assert t.row.tolist() == [0, 1, 2, 4, 4] and t.col.tolist() == [3, 0, 1, 0]


def test_against_reference_parsers(tmp_path):
    lib = ref_parser()
    if lib is None:
        pytest.skip("reference tree not present")
    # A graph without empty rows (a full grid: every vertex has >= 2
    # neighbours), with shuffled in-row order, in both formats.
    g = gdb.gen_grid(17, 19, 0.0, seed=3, wlo=1, whi=50)
    rng = np.random.default_rng(0)
    for v in range(g.n):
        seg = slice(g.row[v], g.row[v + 1])
        p = rng.permutation(g.row[v + 1] - g.row[v])
        g.col[seg], g.w[seg] = g.col[seg][p], g.w[seg][p]
    mf = tmp_path / "g.graph"
    lines = [f"{g.n} {g.m // 2}"]
    for v in range(g.n):
        lines.append(" ".join(str(c + 1) for c in g.col[g.row[v]:g.row[v + 1]]))
    mf.write_text("\n".join(lines) + "\n")
    mine = gdb.load_metis(mf, directed=False)
    r, c, _ = ref_load(lib, mf, 0, 0)
    assert (mine.row == r).all() and (mine.col == c).all()
    # DIMACS with every arc listed (both directions), in a shuffled order.
    df = tmp_path / "g.gr"
    src = np.repeat(np.arange(g.n), g.degrees())
    order = rng.permutation(g.m)
    arcs = [(int(src[k]) + 1, int(g.col[k]) + 1, int(g.w[k])) for k in order]
    df.write_text(f"p sp {g.n} {len(arcs)}\n" + "".join(f"a {a} {b} {w}\n" for a, b, w in arcs))
    for kind, directed, transpose in ((1, 1, False), (1, 0, False), (2, 1, True), (2, 0, True)):
        mine = gdb.load_dimacs(df, directed=bool(directed), transpose=transpose)
        r, c, w = ref_load(lib, df, kind, directed)
        assert (np.diff(mine.row) > 0).all()
        assert (mine.row == r).all() and (mine.col == c).all() and (mine.w == w).all(), (kind, directed)


def test_galois_gr_round_trip(tmp_path):
    g = gdb.gen_rmat(10, 8, seed=2, wlo=1, whi=9)
    f = tmp_path / "g.gr"
    gdb.write_galois_gr(f, g)
    h = gdb.load_galois_gr(f)
    assert h.n == g.n and (h.row == g.row).all() and (h.col == g.col).all() and (h.w == g.w).all()
    # header: version 1, 4-byte edge data, n, m
    hdr = np.fromfile(f, dtype="<u8", count=4)
    assert hdr.tolist() == [1, 4, g.n, g.m]


def test_galois_gr_rejects_bad_version(tmp_path):
    f = tmp_path / "bad.gr"
    np.array([2, 4, 1, 0], dtype="<u8").tofile(f)
    with pytest.raises(gdb.GdbError):
        gdb.load_galois_gr(f)


def test_edges_format(tmp_path):
    f = tmp_path / "g.edges"
    f.write_text("3 4\n0 1 5\n0 2 1\n1 2 2\n2 0 9\n")
    g = gdb.load_edges(f)
    assert g.row.tolist() == [0, 2, 3, 4] and g.col.tolist() == [1, 2, 2, 0] and g.w.tolist() == [5, 1, 2, 9]
