"""Boruvka minimum spanning forest (gdb_mst) against the Kruskal oracle:
weight, edge count and components exact (the forest weight is unique),
including ties, disconnected graphs, isolated vertices and several launch
shapes."""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gdb.Context()
    yield c
    c.close()


GRAPHS = {
    "rmat12": lambda: gdb.gen_rmat(12, 16, seed=12, wlo=1, whi=100000),
    "rmat14_ties": lambda: gdb.gen_rmat(14, 8, seed=3, wlo=1, whi=4),
    "grid200": lambda: gdb.gen_grid(200, 200, 0.05, seed=2, wlo=1, whi=255),
    "gnp_sparse": lambda: gdb.gen_gnp(20000, 1.5, seed=4, wlo=1, whi=1000),
    "equal_weights": lambda: gdb.gen_grid(60, 60, 0.0, seed=5, wlo=3, whi=3),
}


@pytest.mark.parametrize("wgs,groups", [(256, 0), (64, 7), (1024, 32)])
@pytest.mark.parametrize("name", list(GRAPHS))
def test_mst_matches_kruskal(ctx, name, wgs, groups):
    g = GRAPHS[name]()
    want = oracle.mst_kruskal(g.row, g.col, g.w)
    wt, ed, cp, st = gdb.mst(ctx, gdb.DeviceCSR(g), gdb.Opts.make(wgs=wgs, groups=groups))
    print(name, wgs, groups, want, st.rounds, st.kernel_ms)
    assert (wt, ed, cp) == want


def test_mst_small_fixtures(ctx):
    cases = [
        (5, [(0, 1), (1, 2), (1, 3), (3, 4)], [5, 9, 2, 7], (23, 4, 1)),
        (4, [(0, 1), (1, 2), (2, 3), (3, 0)], [4, 1, 8, 3], (8, 3, 1)),
        (6, [(0, 1), (2, 3)], [3, 4], (7, 2, 4)),
        (1, [], [], (0, 0, 1)),
        (3, [(0, 1), (1, 2), (0, 2)], [5, 5, 5], (10, 2, 1)),
    ]
    for n, edges, wts, want in cases:
        row, col, w = csr_from_edges(n, edges, wts)
        g = gdb.CSR(n, row, col, w if len(w) else np.zeros(0, np.int32))
        wt, ed, cp, _ = gdb.mst(ctx, gdb.DeviceCSR(g))
        assert (wt, ed, cp) == want, (n, edges)
