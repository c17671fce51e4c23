"""Pins the CPU oracle before it is trusted as the checker.

Each restatement (orc_*_ref) is fixed by checks that do not retype its
formula:
  color_ref     hand fixtures (incl. the degree-1 quirk and the extra empty
                round) + the longest-path recurrence orc_color_tb on seeded
                R-MAT / G(n,p) graphs;
  mis_ref       hand fixtures + greedy sequential MIS in value order;
  sssp_pull_ref hand fixtures (incl. the n-1 round cap / one-round lag) +
                Dijkstra and Bellman-Ford when the hop diameter is small;
  bfs/sssp_worklist_ref  hand fixtures + queue BFS / Dijkstra + the
                reference's own dverifysolution invariant;
  bc_ref_f32    hand fixtures + Brandes FP64 + the pair-dependency
                definition on tiny graphs + closed forms (path, star);
  block_scan_ref  hand fixture + numpy cumsum.
A planted-mistake check shows the pins bite: a colouring with one vertex
moved a round, a BFS with a level off by one, and a BC value scaled by 1/2
all fail them.
"""
import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle
from conftest import csr_from_edges, grid_csr, load_fixtures, values_of

FX = load_fixtures()


@pytest.mark.parametrize("fx", FX["color"], ids=lambda f: f["name"])
def test_color_fixtures(fx):
    row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    got, rounds = oracle.color_ref(row, col, values_of(fx))
    assert got.tolist() == fx["color"]
    assert rounds == fx["rounds"]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_color_ref_matches_recurrence(seed):
    g = gdb.gen_rmat(10, 8, seed=seed)
    nv = gdb.pannotia_values(g.n)
    a, _ = oracle.color_ref(g.row, g.col, nv)
    b, _ = oracle.color_tb(g.row, g.col, nv)
    assert (a == b).all()
    perm = gdb.permutation(g.n, seed).astype(np.float32) / np.float32(g.n + 1)
    a, _ = oracle.color_ref(g.row, g.col, perm)
    b, _ = oracle.color_tb(g.row, g.col, perm)
    assert (a == b).all()


def test_color_pins_catch_planted_mistake():
    g = gdb.gen_gnp(300, 6.0, seed=9)
    nv = gdb.pannotia_values(g.n)
    a, _ = oracle.color_ref(g.row, g.col, nv)
    bad = a.copy()
    bad[np.argmax(a)] += 1
    b, _ = oracle.color_tb(g.row, g.col, nv)
    assert (a == b).all() and not (bad == b).all()


def test_color_ties_livelock_detected():
    row, col, _ = csr_from_edges(3, [(0, 1), (1, 2), (0, 2)])
    _, rounds = oracle.color_ref(row, col, np.array([0.5, 0.5, 0.5], np.float32))
    assert rounds == -1


@pytest.mark.parametrize("fx", FX["mis"], ids=lambda f: f["name"])
def test_mis_fixtures(fx):
    row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    got, _ = oracle.mis_ref(row, col, values_of(fx))
    assert got.tolist() == fx["s"]


@pytest.mark.parametrize("seed", [1, 5])
def test_mis_ref_matches_greedy(seed):
    g = gdb.gen_rmat(11, 8, seed=seed)
    for nv in (gdb.pannotia_values(g.n),
               gdb.permutation(g.n, seed).astype(np.float32) / np.float32(g.n + 1)):
        a, _ = oracle.mis_ref(g.row, g.col, nv)
        b = oracle.mis_tb(g.row, g.col, nv)
        assert (a == b).all()
        # independence and maximality, from the definition
        ins = a == 2
        src = np.repeat(np.arange(g.n), np.diff(g.row))
        assert not (ins[src] & ins[g.col]).any()
        covered = ins.copy()
        covered[g.col[ins[src]]] = True
        assert covered.all()


@pytest.mark.parametrize("fx", FX["sssp_pull"], ids=lambda f: f["name"])
def test_sssp_pull_fixtures(fx):
    row, col, w = csr_from_edges(fx["n"], fx["edges"], fx["w"])
    got, _ = oracle.sssp_pull_ref(row, col, w, 0)  # symmetric: in-edges == out-edges
    assert got.tolist() == fx["dist"]


def test_sssp_pull_matches_dijkstra_and_bellman_ford():
    g = gdb.gen_rmat(11, 8, seed=4, wlo=1, whi=100)
    gt = g.transpose()
    x, rounds = oracle.sssp_pull_ref(gt.row, gt.col, gt.w, 0)
    assert rounds < g.n - 1
    d = oracle.dijkstra_tb(g.row, g.col, g.w, 0, oracle.PANNOTIA_INF)
    bf = oracle.bellman_ford_tb(g.row, g.col, g.w, 0, oracle.PANNOTIA_INF)
    assert (d == bf).all()
    assert (x == d).all()


@pytest.mark.parametrize("fx", FX["bfs"], ids=lambda f: f["name"])
def test_bfs_fixtures(fx):
    if "rows" in fx:
        row, col, _ = grid_csr(fx["rows"], fx["cols"])
    else:
        row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    got, _ = oracle.bfs_worklist_ref(row, col, 0)
    assert got.tolist() == fx["dist"]


def test_bfs_worklist_matches_queue_bfs_and_grid_closed_form():
    g = gdb.gen_grid(40, 50, 0.0, seed=2)
    got, iters = oracle.bfs_worklist_ref(g.row, g.col, 0)
    r, c = np.divmod(np.arange(g.n), 50)
    assert (got == r + c).all()
    assert iters == 40 + 50 - 2 + 1
    g = gdb.gen_grid(60, 60, 0.05, seed=2)
    got, _ = oracle.bfs_worklist_ref(g.row, g.col, 0)
    tb = oracle.bfs_tb(g.row, g.col, 0)
    assert (got == tb).all()
    assert oracle.verify_distances_ref(g.row, g.col, None, got) == 0
    bad = got.copy()
    bad[np.flatnonzero(got == 5)[0]] = 7
    assert oracle.verify_distances_ref(g.row, g.col, None, bad) > 0


@pytest.mark.parametrize("fx", FX["sssp_worklist"], ids=lambda f: f["name"])
def test_sssp_worklist_fixtures(fx):
    row, col, w = csr_from_edges(fx["n"], fx["edges"], fx["w"])
    got, _ = oracle.sssp_worklist_ref(row, col, w, 0)
    assert got.tolist() == fx["dist"]


def test_sssp_worklist_matches_dijkstra():
    g = gdb.gen_grid(50, 50, 0.05, seed=2, wlo=1, whi=255)
    got, _ = oracle.sssp_worklist_ref(g.row, g.col, g.w, 0)
    d = oracle.dijkstra_tb(g.row, g.col, g.w, 0, oracle.LONESTAR_INF)
    assert (got.astype(np.int64) == d).all()
    assert oracle.verify_distances_ref(g.row, g.col, g.w, got) == 0


@pytest.mark.parametrize("fx", FX["bc"], ids=lambda f: f["name"])
def test_bc_fixtures(fx):
    row, col, _ = csr_from_edges(fx["n"], fx["edges"])
    src = np.array(fx["sources"], np.int32)
    np.testing.assert_allclose(oracle.bc_brandes_tb(row, col, src), fx["bc"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(oracle.bc_pairs_tb(row, col, src), fx["bc"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(oracle.bc_ref_f32(row, col, row, col, src), fx["bc"], rtol=1e-6, atol=1e-6)


def test_bc_brandes_matches_pair_definition():
    g = gdb.gen_gnp(60, 4.0, seed=11)
    src = np.arange(0, 60, 3, dtype=np.int32)
    a = oracle.bc_brandes_tb(g.row, g.col, src)
    b = oracle.bc_pairs_tb(g.row, g.col, src)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-9)
    assert not np.allclose(a * 0.5, b)  # a dropped/added 1/2 factor would fail
    f = oracle.bc_ref_f32(g.row, g.col, g.row, g.col, src)
    np.testing.assert_allclose(f, a, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_worklist_threaded_matches_queue_bfs_and_dijkstra(threads):
    # The threaded CPU baseline (persistent threads, CPU XF barrier) against
    # queue BFS and Dijkstra, and its BFS iteration count against the
    # restatement's (levels + 1 on a connected graph).
    for g in (gdb.gen_grid(70, 90, 0.05, seed=2, wlo=1, whi=255), gdb.gen_rmat(12, 16, seed=6, wlo=1, whi=255)):
        for src in (0, 17):
            got, it = oracle.worklist_mt(g.row, g.col, None, src, threads=threads)
            assert (got == oracle.bfs_tb(g.row, g.col, src)).all()
            want, it_ref = oracle.bfs_worklist_ref(g.row, g.col, src)
            assert (got == want).all() and it == it_ref
            got_s, _ = oracle.worklist_mt(g.row, g.col, g.w, src, threads=threads)
            assert (got_s.astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, src)).all()


def test_worklist_threaded_planted_mistake_is_caught():
    # A distance one too large at a reachable vertex fails the Dijkstra pin.
    g = gdb.gen_grid(30, 30, 0.0, seed=1, wlo=1, whi=9)
    got, _ = oracle.worklist_mt(g.row, g.col, g.w, 0, threads=4)
    bad = got.copy()
    bad[400] += 1
    assert not (bad.astype(np.int64) == oracle.dijkstra_tb(g.row, g.col, g.w, 0)).all()


def test_bc_brandes_threaded_matches_pair_definition():
    # The threaded form (sources dealt to threads, accumulators summed in
    # thread order) against the pair-dependency definition and the
    # single-thread form, with more threads than sources and uneven shares.
    g = gdb.gen_gnp(60, 4.0, seed=12)
    src = np.arange(1, 60, 4, dtype=np.int32)
    want = oracle.bc_pairs_tb(g.row, g.col, src)
    for t in (1, 3, 7, 64):
        np.testing.assert_allclose(oracle.bc_brandes_mt(g.row, g.col, src, threads=t), want, rtol=1e-12, atol=1e-9)
    r = gdb.gen_rmat(11, 16, seed=4)
    s = gdb.permutation(r.n, 4)[:21]
    np.testing.assert_allclose(oracle.bc_brandes_mt(r.row, r.col, s, threads=5),
                               oracle.bc_brandes_tb(r.row, r.col, s), rtol=1e-12, atol=1e-9)


def test_bc_path_closed_form():
    n = 9
    row, col, _ = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    src = np.arange(n, dtype=np.int32)
    v = np.arange(n)
    np.testing.assert_allclose(oracle.bc_brandes_tb(row, col, src), 2 * v * (n - 1 - v), atol=1e-9)


@pytest.mark.parametrize("fx", FX["block_scan"], ids=lambda f: f["name"])
def test_block_scan_fixture(fx):
    out, total = oracle.block_scan_ref(fx["in"])
    assert out.tolist() == fx["out"] and total == fx["total"]


def test_block_scan_matches_cumsum():
    rng = np.random.default_rng(0)
    for n in (1, 2, 64, 256, 1024):
        v = rng.integers(0, 50, n).astype(np.int32)
        out, total = oracle.block_scan_ref(v)
        assert (out == np.concatenate([[0], np.cumsum(v)[:-1]])).all() and total == v.sum()
