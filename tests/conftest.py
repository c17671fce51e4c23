import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_fixtures():
    return json.loads((GOLDEN / "hand_fixtures.json").read_text())


def csr_from_edges(n, edges, weights=None):
    """Symmetric CSR (both directions, rows sorted) from an undirected edge list."""
    adj = [[] for _ in range(n)]
    for k, (u, v) in enumerate(edges):
        w = 1 if weights is None else weights[k]
        adj[u].append((v, w))
        adj[v].append((u, w))
    row = np.zeros(n + 1, np.int32)
    col, wt = [], []
    for i in range(n):
        for v, w in sorted(adj[i]):
            col.append(v)
            wt.append(w)
        row[i + 1] = len(col)
    return row, np.array(col, np.int32), np.array(wt, np.int32)


def grid_csr(rows, cols):
    edges = []
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                edges.append((v, v + 1))
            if r + 1 < rows:
                edges.append((v, v + cols))
    return csr_from_edges(rows * cols, edges)


def values_of(fx):
    n = fx["n"]
    if "values" in fx:
        return np.array(fx["values"], np.float32)
    return (np.arange(n, dtype=np.float32) / np.float32(n + 1)).astype(np.float32)
