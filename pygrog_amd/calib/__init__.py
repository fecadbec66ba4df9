"""Calibration: GRAPPA operator table and GROG interpolation (reference
src/pygrog/calib/__init__.py)."""

from ._acr import extract_acr
from ._grappa import KernelTable, KernelTableResult, build_table, train_grappa
from ._grog import GrogInterpolator, GrogPlan
from ._batch import GrogFrameBatch

__all__ = ["GrogFrameBatch", "GrogInterpolator", "GrogPlan", "KernelTable", "KernelTableResult", "build_table",
           "extract_acr", "train_grappa"]
