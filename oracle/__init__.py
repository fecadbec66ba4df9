"""Oracle of the GROG gridding path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker: nothing in
``pygrog_amd`` imports it.

Contents
  * ``ops``:   numpy wrappers of ``grog_oracle.c``, the plain C restatement of
               the reference's CPU kernels (built by ``make -C oracle``).
  * ``plan``:  numpy float32 restatement of the reference's plan construction
               and shift quantisation.
  * ``table``: FP64 (complex128) GRAPPA operator table with numpy/scipy.
  * ``grid``:  the reference's ``interpolate(grid=True)`` flow on the C ops,
               and an FP64 gridding oracle.
  * ``rows``:  what the grouped path's lossy sample-row formats (24-bit,
               block floating point) must hold, in numpy.
  * ``ref_ext()``: the reference's own CPU ops (``csrc/torch``) compiled in
               place into ``oracle/_ref`` by ``make -C oracle ref``, when built.
"""

from __future__ import annotations

import ctypes
import pathlib
import subprocess
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "_build" / "libgrog_oracle.so"
REF_DIR = HERE / "_ref"

_lib = None


def build() -> None:
    """Compile the C restatement (and the reference ops when the reference
    sources are present)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if pathlib.Path("/root/reference/csrc/torch/module.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", "-j4"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        _lib = ctypes.CDLL(str(LIB))
    return _lib


def ref_ext():
    """The reference's compiled CPU ops module, or None if not built."""
    if not any(REF_DIR.glob("_pygrog_torch_ref*.so")):
        return None
    import torch  # noqa: F401  (loads libtorch first)

    if str(REF_DIR) not in sys.path:
        sys.path.insert(0, str(REF_DIR))
    import _pygrog_torch_ref  # type: ignore

    return _pygrog_torch_ref


from . import grid, ops, plan, table  # noqa: E402,F401

__all__ = ["build", "grid", "lib", "ops", "plan", "ref_ext", "table"]
_ = np
