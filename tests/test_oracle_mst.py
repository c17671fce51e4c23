"""Pins the minimum-spanning-forest oracle (orc_mst_kruskal) before it checks
the Boruvka kernel: Kruskal == Prim (two textbook algorithms) on seeded
graphs, closed forms (a tree, a cycle, a unit-weight grid, isolated
vertices) and brute force over every spanning tree of small complete graphs.
A planted mistake (a maximum spanning tree) fails the pins."""
import itertools

import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle
from conftest import csr_from_edges


@pytest.mark.parametrize("kind", ["rmat", "gnp", "grid", "ties"])
def test_kruskal_equals_prim(kind):
    if kind == "rmat":
        g = gdb.gen_rmat(11, 8, seed=5, wlo=1, whi=1000)
    elif kind == "gnp":
        g = gdb.gen_gnp(3000, 3.0, seed=6, wlo=1, whi=50)
    elif kind == "grid":
        g = gdb.gen_grid(40, 30, 0.1, seed=7, wlo=1, whi=255)
    else:
        g = gdb.gen_grid(20, 20, 0.0, seed=8, wlo=7, whi=7)
    a = oracle.mst_kruskal(g.row, g.col, g.w)
    b = oracle.mst_prim_tb(g.row, g.col, g.w)
    assert a == b
    assert a[1] + a[2] == g.n


def test_closed_forms():
    # A tree: every edge is in the forest.
    edges = [(0, 1), (1, 2), (1, 3), (3, 4)]
    wts = [5, 9, 2, 7]
    row, col, w = csr_from_edges(5, edges, wts)
    assert oracle.mst_kruskal(row, col, w) == (23, 4, 1)
    # A cycle of distinct weights: all but the heaviest edge.
    edges = [(0, 1), (1, 2), (2, 3), (3, 0)]
    row, col, w = csr_from_edges(4, edges, [4, 1, 8, 3])
    assert oracle.mst_kruskal(row, col, w) == (8, 3, 1)
    # Unit-weight grid: n - 1 edges of weight 1.
    g = gdb.gen_grid(12, 9, 0.0, seed=1, wlo=1, whi=1)
    assert oracle.mst_kruskal(g.row, g.col, g.w) == (g.n - 1, g.n - 1, 1)
    # Isolated vertices are components of their own.
    row, col, w = csr_from_edges(6, [(0, 1), (2, 3)], [3, 4])
    assert oracle.mst_kruskal(row, col, w) == (7, 2, 4)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_brute_force_small_complete_graphs(seed):
    n = 6
    rng = np.random.default_rng(seed)
    pairs = list(itertools.combinations(range(n), 2))
    wts = rng.integers(1, 6, len(pairs)).tolist()  # ties on purpose
    row, col, w = csr_from_edges(n, pairs, wts)
    best = None
    for sub in itertools.combinations(range(len(pairs)), n - 1):
        par = list(range(n))

        def f(x):
            while par[x] != x:
                x = par[x]
            return x

        ok = True
        for i in sub:
            a, b = f(pairs[i][0]), f(pairs[i][1])
            if a == b:
                ok = False
                break
            par[a] = b
        if ok:
            tot = sum(wts[i] for i in sub)
            best = tot if best is None else min(best, tot)
    assert oracle.mst_kruskal(row, col, w) == (best, n - 1, 1)


def test_planted_mistake_fails():
    # A maximum spanning tree (Kruskal on negated weights) differs.
    g = gdb.gen_gnp(400, 4.0, seed=9, wlo=1, whi=100)
    neg = (101 - g.w).astype(np.uint32)
    wmax, ed, _ = oracle.mst_kruskal(g.row, g.col, neg)
    wrong = 101 * ed - wmax
    assert wrong != oracle.mst_kruskal(g.row, g.col, g.w)[0]
