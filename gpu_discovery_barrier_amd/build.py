"""Builds libgdb_amd.so (HIP, gfx950) and the command-line tools in-tree.

Usage: python -m gpu_discovery_barrier_amd.build [--jobs N] [--verbose]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libgdb_amd.so"
BIN = PKG / "bin"

ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    # Packed FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROOT / 'include'}", f"-I{CSRC}"]

# Command-line tools: (executable, source, extra defines).  The names are the
# reference's executables (pannotia/CMakeLists.txt:38-89,
# lonestar_gpu_opencl/CMakeLists.txt:32-83, occupancy_tests/CMakeLists.txt).
TOOLS = [
    ("color", "cli/pannotia_app.cpp", ["-DGDB_APP_COLOR", "-DGDB_APP_RELAUNCH"]),
    ("color-gb", "cli/pannotia_app.cpp", ["-DGDB_APP_COLOR"]),
    ("mis", "cli/pannotia_app.cpp", ["-DGDB_APP_MIS", "-DGDB_APP_RELAUNCH"]),
    ("mis-gb", "cli/pannotia_app.cpp", ["-DGDB_APP_MIS"]),
    ("sssp", "cli/pannotia_app.cpp", ["-DGDB_APP_SSSP", "-DGDB_APP_RELAUNCH"]),
    ("sssp-gb", "cli/pannotia_app.cpp", ["-DGDB_APP_SSSP"]),
    ("bc", "cli/pannotia_app.cpp", ["-DGDB_APP_BC", "-DGDB_APP_RELAUNCH"]),
    ("bc-gb", "cli/pannotia_app.cpp", ["-DGDB_APP_BC"]),
    ("bfs-port", "cli/lonestar_app.cpp", ["-DGDB_APP_BFS"]),
    ("sssp-port", "cli/lonestar_app.cpp", ["-DGDB_APP_SSSP"]),
    ("mst-port", "cli/lonestar_app.cpp", ["-DGDB_APP_MST"]),
    ("occupancy_test", "cli/occupancy_tools.cpp", ["-DGDB_TOOL_OCCUPANCY"]),
    ("time_prot", "cli/occupancy_tools.cpp", ["-DGDB_TOOL_TIME_PROT"]),
    ("device_query", "cli/occupancy_tools.cpp", ["-DGDB_TOOL_DEVICE_QUERY"]),
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libgdb_amd.so")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... ({res.returncode})")
    elif verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)


def sources() -> tuple[list[Path], list[Path]]:
    hip = sorted(CSRC.glob("*.hip"))
    cpp = sorted(CSRC.glob("*.cpp"))
    return hip, cpp


def build(jobs: int | None = None, verbose: bool = False, tools: bool = True) -> Path:
    """Compiles every HIP source for gfx950 and links libgdb_amd.so."""
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    hip, cpp = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    cmds = []
    objs = []
    for src in hip:
        obj = OBJ / (src.stem + ".hip.o")
        objs.append(obj)
        cmds.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in cpp:
        obj = OBJ / (src.stem + ".cpp.o")
        objs.append(obj)
        cmds.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    if tools:
        build_tools(jobs, verbose)
    return LIB


def build_tools(jobs: int | None = None, verbose: bool = False) -> None:
    BIN.mkdir(exist_ok=True)
    cmds = []
    for name, src, defs in TOOLS:
        path = CSRC / src
        if not path.exists():
            continue
        cmds.append(["g++", *CXX_FLAGS, *defs, str(path), str(CSRC / "cli" / "cli_common.cpp"),
                     "-o", str(BIN / name), f"-L{PKG}", "-lgdb_amd", f"-Wl,-rpath,{PKG}",
                     "-Wl,-rpath,$ORIGIN/.."])
    with cf.ThreadPoolExecutor(max_workers=jobs or 8) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-tools", action="store_true")
    args = ap.parse_args()
    lib = build(args.jobs, args.verbose, tools=not args.no_tools)
    print(f"built {lib}")


if __name__ == "__main__":
    main()
