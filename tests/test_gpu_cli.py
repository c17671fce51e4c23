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
