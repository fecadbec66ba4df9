"""pygrog_amd: MI355X-native GROG gridding (drop-in for PyGROG's gridding path).

Public names follow the reference package (src/pygrog/__init__.py):
``calib`` (GrogInterpolator, GrogPlan, KernelTable), ``operator`` (SparseFFT,
MaskedFFT, MaskedFFTPlan, scatter_add, gather), ``cg`` / ``lsmr`` / ``solve``
and ``PolynomialPreconditioner``.  Compute runs in HIP kernels of
``_lib/libpygrog_hip.so`` (built by ``pygrog_amd.csrc.build``).
"""

__version__ = "0.1.0"

from . import calib  # noqa: E402,F401
from . import operator  # noqa: E402,F401
from ._solve import PolynomialPreconditioner, SolveMixin, cg, lsmr, solve  # noqa: E402,F401

__all__ = ["PolynomialPreconditioner", "SolveMixin", "calib", "cg", "lsmr", "operator", "solve"]
