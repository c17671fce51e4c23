"""The rebuilt command-line tools keep the reference's usage, stdout strings
and output files (callers such as the reference's run_suite.py parse them):
color-gb on a METIS file (color_gb.cpp: color_gb.out "%d: %d" 1-based),
bfs-port / sssp-port / mst-port on a Galois .gr file (bfs_port.cpp,
sssp_port.cpp, mst_port.cpp stdout).  Results are checked against the
oracle."""
import re
import subprocess

import numpy as np
import pytest

import gpu_discovery_barrier_amd as gdb
import oracle

pytestmark = pytest.mark.gpu

BIN = gdb.api.PKG / "bin"


def run(args, cwd):
    p = subprocess.run([str(BIN / args[0]), *map(str, args[1:])], cwd=cwd, capture_output=True, text=True,
                       timeout=120)
    print(p.stdout[-2000:], p.stderr[-2000:])
    return p


def write_metis(path, g):
    """METIS: "n m" then each vertex's 1-based neighbours (m undirected edges)."""
    with open(path, "w") as f:
        f.write(f"{g.n} {g.m // 2}\n")
        for v in range(g.n):
            f.write(" ".join(str(int(u) + 1) for u in g.col[g.row[v]:g.row[v + 1]]) + "\n")


@pytest.fixture(scope="module")
def grid():
    return gdb.gen_grid(48, 40, 0.0, seed=3, wlo=1, whi=255)


def test_color_gb_cli(tmp_path, grid):
    f = tmp_path / "g.graph"
    write_metis(f, grid)
    p = run(["color-gb", f, 1, 256], tmp_path)
    assert p.returncode == 0
    assert re.search(r"kernel ran with a total of \d+ workgroups", p.stdout)
    assert re.search(r"kernel time = [\d.]+ ms", p.stdout)
    got = np.zeros(grid.n, np.int64)
    for line in (tmp_path / "color_gb.out").read_text().split("\n"):
        if line.strip():
            a, b = line.split(":")
            got[int(a) - 1] = int(b)
    want, _ = oracle.color_ref(grid.row, grid.col, gdb.pannotia_values(grid.n))
    assert (got == want).all()


@pytest.mark.parametrize("exe", ["bfs-port", "sssp-port"])
def test_lonestar_port_cli(tmp_path, grid, exe):
    f = tmp_path / "g.gr"
    gdb.write_galois_gr(str(f), grid)
    p = run([exe, f, 256], tmp_path)
    assert p.returncode == 0
    assert re.search(r"\tapp runtime = [\d.]+ ms\.", p.stdout)
    assert re.search(r"\tnumber of participating groups = \d+", p.stdout)
    assert "\tno of errors = 0." in p.stdout


def test_mst_port_cli(tmp_path):
    g = gdb.gen_rmat(11, 8, seed=21, wlo=1, whi=1000)
    f = tmp_path / "g.sym.gr"
    gdb.write_galois_gr(str(f), g)
    p = run(["mst-port", f, 128], tmp_path)
    assert p.returncode == 0
    m = re.search(r"result: weight: (\d+), components: (\d+), edges: (\d+)", p.stdout)
    assert m
    w, edges, comps = oracle.mst_kruskal(g.row, g.col, g.w)
    assert (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (w & 0xFFFFFFFF, comps, edges)


def write_dimacs(path, g):
    """DIMACS shortest-path format: "p sp n m", then "a u v w" (1-based), every arc."""
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    with open(path, "w") as f:
        f.write(f"c grid\np sp {g.n} {g.m}\n")
        for u, v, w in zip(src, g.col, g.w):
            f.write(f"a {u + 1} {v + 1} {w}\n")


def read_pairs(path, n, conv=int):
    got = np.zeros(n, np.int64 if conv is int else np.float64)
    for line in path.read_text().split("\n"):
        if line.strip():
            a, b = line.split(":")
            got[int(a) - 1] = conv(b)
    return got


def test_color_relaunch_cli(tmp_path, grid):
    # color.cpp: relaunch loop; "total number of colors used" is rounds + 1.
    f = tmp_path / "g.graph"
    write_metis(f, grid)
    p = run(["color", f, 1, 128], tmp_path)
    assert p.returncode == 0
    want, rounds = oracle.color_ref(grid.row, grid.col, gdb.pannotia_values(grid.n))
    assert re.search(r"total number of colors used: \d+", p.stdout)
    assert re.search(r"kernel time = [\d.]+ ms", p.stdout)
    assert (read_pairs(tmp_path / "color.out", grid.n) == want).all()


@pytest.mark.parametrize("exe,out", [("mis", "mis.out"), ("mis-gb", "mis_gb.out")])
def test_mis_cli(tmp_path, grid, exe, out):
    f = tmp_path / "g.graph"
    write_metis(f, grid)
    p = run([exe, f, 1, 256], tmp_path)
    assert p.returncode == 0
    if exe.endswith("-gb"):
        assert re.search(r"kernel ran with a total of \d+ workgroups", p.stdout)
    got = np.array([int(x) for x in (tmp_path / out).read_text().split()], np.int64)
    want, _ = oracle.mis_ref(grid.row, grid.col, gdb.pannotia_values(grid.n))
    assert (got == want).all()


@pytest.mark.parametrize("exe,out", [("sssp", "sssp.out"), ("sssp-gb", "sssp_gb.out")])
def test_sssp_cli(tmp_path, exe, out):
    # DIMACS is read transposed (sssp_gb.cpp:124-133) and relaxed over in-edges
    # from vertex 0; a directed graph makes the transpose matter.
    g = gdb.gen_rmat(10, 8, seed=31, wlo=1, whi=100)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    keep = src < g.col
    row = np.zeros(g.n + 1, np.int64)
    np.add.at(row, src[keep] + 1, 1)
    dg = gdb.CSR(g.n, np.cumsum(row).astype(np.int32), g.col[keep].astype(np.int32), g.w[keep].astype(np.int32))
    f = tmp_path / "g.gr"
    write_dimacs(f, dg)
    p = run([exe, f, 0, 128], tmp_path)
    assert p.returncode == 0
    assert re.search(r"kernel time = [\d.]+ ms", p.stdout)
    gt = dg.transpose()
    want, _ = oracle.sssp_pull_ref(gt.row, gt.col, gt.w, 0)
    assert (read_pairs(tmp_path / out, g.n) == want).all()


@pytest.mark.parametrize("exe,out", [("bc", "bc.out"), ("bc-gb", "bc_gb.out")])
def test_bc_cli(tmp_path, grid, exe, out):
    # Directed DIMACS (bc_gb.cpp:105,131); the first 24 vertices as sources
    # (the reference loops over all n); "%f" output checked to its printed
    # precision against FP64 Brandes.
    f = tmp_path / "g.gr"
    write_dimacs(f, grid)
    env = {"GDB_BC_SOURCES": "24"}
    pr = subprocess.run([str(BIN / exe), str(f), "0", "256"], cwd=tmp_path, capture_output=True, text=True,
                        timeout=120, env={**__import__("os").environ, **env})
    print(pr.stdout[-2000:], pr.stderr[-2000:])
    assert pr.returncode == 0
    got = np.array([float(x) for x in (tmp_path / out).read_text().split()])
    want = oracle.bc_brandes_tb(grid.row, grid.col, np.arange(24, dtype=np.int32))
    assert np.all(np.abs(got - want) <= np.maximum(1e-4 * np.abs(want), 1e-6) + 5e-7)


def test_occupancy_test_cli(tmp_path):
    bound = gdb.max_resident_groups(256, 4)
    p = run(["occupancy_test", 1000, 256, 0, 1, 2], tmp_path)
    assert p.returncode == 0
    m = re.search(r"kernel ran with a total of (\d+) workgroups", p.stdout)
    assert m and 1 <= int(m.group(1)) <= min(1000, bound)
    assert "protocol enabled: 1" in p.stdout
    # Skip mode at the bound completes with every group; one more is refused.
    p = run(["occupancy_test", bound, 256, 0, 0, 2], tmp_path)
    assert p.returncode == 0
    assert f"kernel ran with a total of {bound} workgroups" in p.stdout
    p = run(["occupancy_test", bound + 1, 256, 0, 0, 2], tmp_path)
    assert p.returncode == 2 and "refused" in p.stdout


@pytest.mark.parametrize("ticket", [0, 1, 2])
def test_time_prot_cli(tmp_path, ticket):
    p = run(["time_prot", 1000, 256, 0, ticket], tmp_path)
    assert p.returncode == 0
    assert re.search(r"kernel ran with a total of [1-9]\d* workgroups", p.stdout)
    m = re.search(r"kernel time: ([\d.]+) ms", p.stdout)
    assert m and float(m.group(1)) > 0


def test_device_query_cli(tmp_path):
    p = run(["device_query"], tmp_path)
    assert p.returncode == 0
    assert "gfx950" in p.stdout
    m = re.search(r"compute units: (\d+)", p.stdout)
    assert m and int(m.group(1)) >= 1
