"""Iterative solvers (CG, LSMR), polynomial preconditioning and the
``op.solve`` mixin (reference src/pygrog/_solve/)."""

from ._mixin import SolveMixin
from ._polynomial import PolynomialPreconditioner
from ._solvers import cg, lsmr, solve

__all__ = ["PolynomialPreconditioner", "SolveMixin", "cg", "lsmr", "solve"]
