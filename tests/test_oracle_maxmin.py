"""Pins of the max-min colouring oracle (orc_color_maxmin_ref, no reference
counterpart: the two-sided variant of the reference's max rule,
color_kernel.cl:229-237).  None of these retypes the rule: properness and
colour-class independence on random graphs, closed forms on complete graphs,
monotone paths and stars, and planted mistakes the pins reject."""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle
from conftest import csr_from_edges


def perm_values(n, seed):
    return (gdb.permutation(n, seed).astype(np.float64) / (n + 1)).astype(np.float32)


def proper(row, col, color):
    src = np.repeat(np.arange(row.size - 1), np.diff(row))
    ne = src != col
    return bool((color[src[ne]] != color[col[ne]]).all())


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_maxmin_is_a_proper_colouring(seed):
    for g in (gdb.gen_rmat(11, 16, seed=seed), gdb.gen_gnp(3000, 8.0, seed=seed), gdb.gen_grid(40, 50, 0.1, seed=seed)):
        nv = perm_values(g.n, seed)
        c, r = oracle.color_maxmin_ref(g.row, g.col, nv)
        assert r > 0 and (c >= 1).all() and c.max() <= 2 * r
        assert proper(g.row, g.col, c)


def test_maxmin_complete_graph_closed_form():
    # K_n with distinct values: round r colours the r-th largest (2r-1) and
    # the r-th smallest (2r); with n odd the middle one is both, and takes 2r-1.
    for n in (1, 2, 5, 8):
        edges = [(i, j) for i in range(n) for j in range(i + 1, n)]
        row, col, _ = csr_from_edges(n, edges)
        nv = perm_values(n, 7)
        c, r = oracle.color_maxmin_ref(row, col, nv)
        rank = np.argsort(np.argsort(-nv))  # 0 = largest
        want = np.where(rank < (n + 1) // 2, 2 * rank + 1, 2 * (n - 1 - rank) + 2)
        assert (c == want).all() and r == (n + 1) // 2


def test_maxmin_monotone_path_closed_form():
    # Path 0-1-...-(n-1) with increasing values: vertex i >= n/2 is a maximum in
    # round n-i, vertex i < n/2 a minimum in round i+1; the middle of an odd
    # path is both and takes the odd colour.
    for n in (2, 7, 10):
        row, col, _ = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])
        nv = gdb.pannotia_values(n)
        c, r = oracle.color_maxmin_ref(row, col, nv)
        i = np.arange(n)
        want = np.where(i >= n // 2, 2 * (n - 1 - i) + 1, 2 * i + 2)
        assert (c == want).all()


def test_maxmin_star_and_self_loop():
    # Star with the centre lowest: every leaf is a maximum (1), the centre a
    # minimum (2).  A self-loop does not make a vertex its own neighbour.
    row, col, _ = csr_from_edges(5, [(0, i) for i in range(1, 5)] + [(2, 2)])
    c, r = oracle.color_maxmin_ref(row, col, gdb.pannotia_values(5))
    assert c.tolist() == [2, 1, 1, 1, 1] and r == 1


def test_maxmin_ties_livelock_is_reported():
    row, col, _ = csr_from_edges(3, [(0, 1), (1, 2)])
    c, r = oracle.color_maxmin_ref(row, col, np.array([0.5, 0.5, 0.5], np.float32))
    assert r == -1


def test_maxmin_pins_reject_planted_mistakes():
    g = gdb.gen_rmat(10, 16, seed=5)
    nv = perm_values(g.n, 5)
    c, _ = oracle.color_maxmin_ref(g.row, g.col, nv)
    # A colour copied onto a neighbour breaks properness.
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    e = int(np.flatnonzero(src != g.col)[0])
    bad = c.copy()
    bad[g.col[e]] = bad[src[e]]
    assert not proper(g.row, g.col, bad)
    # The max-only rule's colouring differs (it is not this rule).
    cm, _ = oracle.color_ref(g.row, g.col, nv)
    assert not (cm == c).all()
