"""Build libpygrog_hip.so for gfx950 with hipcc (no torch extension machinery).

Each ``.hip`` file is compiled to an object in parallel, then linked into
``pygrog_amd/_lib/libpygrog_hip.so``.  The plan compiler is built with
``-ffp-contract=off`` so its float32 arithmetic is rounded op by op (the
bit-exact target/shift/table-entry contract), and so are the reference's
seven ops (each product and sum rounded as in the reference's CPU loops);
every file is built with packed FP32 instructions disabled.

Usage: ``python -m pygrog_amd.csrc.build [--jobs N]``.
"""

from __future__ import annotations

import argparse
import os
import pathlib
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = pathlib.Path(__file__).resolve().parent
PKG = HERE.parent
ROOT = PKG.parent
LIB_DIR = PKG / "_lib"
OBJ_DIR = LIB_DIR / "obj"
LIB_PATH = LIB_DIR / "libpygrog_hip.so"

ARCH = "gfx950"
SOURCES = {
    "ops_ref.hip": ["-ffp-contract=off"],
    "plan.hip": ["-ffp-contract=off"],
    "plan_compile.hip": [],
    "grid.hip": [],
    "grid_mfma.hip": [],
    "sense.hip": [],
}
COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _hipcc() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return cand


def _compile(src: str, extra: list[str]) -> pathlib.Path:
    obj = OBJ_DIR / (pathlib.Path(src).stem + ".o")
    cmd = [_hipcc(), *COMMON_FLAGS, *extra, "-c", str(HERE / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> pathlib.Path:
    """Compile every HIP source and link the shared library; return its path."""
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)]
    cmd += [str(o) for o in objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    build(args.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
