"""FP64 GRAPPA operator table with numpy/scipy (TEST INFRASTRUCTURE ONLY).

Same algorithm as the reference (src/pygrog/calib/_grappa.py): Frobenius
normalisation (:284), ridge least squares on the augmented system solved by
LAPACK gelsd (scipy.linalg.lstsq; the reference calls the same driver through
torch, :16-59), operators G = X^T (:343-344, 365-367), fractional powers
expm(p * logm(G)) with logm from the eigendecomposition (:62-72, 184-208),
exponents p_i = (i - (n-1)/2) * stepsize in float32 (:100-103) and the
outer-product table order (:131-181), all in complex128.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from .plan import quant_params


def ridge(A: np.ndarray, B: np.ndarray, lamda: float) -> np.ndarray:
    A = np.asarray(A, dtype=np.complex128)
    B = np.asarray(B, dtype=np.complex128)
    n = A.shape[-1]
    if lamda > 0:
        A = np.concatenate([A, np.sqrt(lamda) * np.eye(n)], axis=0)
        B = np.concatenate([B, np.zeros((n, B.shape[-1]))], axis=0)
    return sla.lstsq(A, B, lapack_driver="gelsd")[0]


def unit_operators(calib: np.ndarray, lamda: float):
    c = np.asarray(calib, dtype=np.complex128)
    c = c / np.linalg.norm(c)
    c = np.moveaxis(c, 0, -1)
    nc = c.shape[-1]
    if c.ndim == 3:
        Gy = ridge(c[:, :-1].reshape(-1, nc), c[:, 1:].reshape(-1, nc), lamda).T
        Gx = ridge(c[:-1].reshape(-1, nc), c[1:].reshape(-1, nc), lamda).T
        return {"x": Gx, "y": Gy, "z": None}
    Gz = ridge(c[:-1].reshape(-1, nc), c[1:].reshape(-1, nc), lamda).T
    Gy = ridge(c[:, :-1].reshape(-1, nc), c[:, 1:].reshape(-1, nc), lamda).T
    Gx = ridge(c[:, :, :-1].reshape(-1, nc), c[:, :, 1:].reshape(-1, nc), lamda).T
    return {"x": Gx, "y": Gy, "z": Gz}


def radial_unit_operators(calib: np.ndarray, coords: np.ndarray, lamda: float = 0.0):
    """Radial self-calibration (reference _grappa.py:290-323): per spoke a
    least-squares operator from each readout sample to the next
    (T = S G, row vectors), log of each, a least-squares fit
    log G_spoke = dx log Gx + dy log Gy over the spokes with the logs
    flattened column-major and unflattened row-major, then expm.  calib
    (coils, spokes, readout), coords (spokes, readout, 2)."""
    c = np.asarray(calib, dtype=np.complex128)
    c = c / np.linalg.norm(c)
    c = np.moveaxis(c, 0, -1)  # (nr, ns, nc)
    nr, ns, nc = c.shape
    lth = np.empty((nr, nc * nc), np.complex128)
    for r in range(nr):
        G = ridge(c[r, :-1], c[r, 1:], lamda)
        lth[r] = logm_eig(G).T.reshape(-1)
    k = np.asarray(coords, np.float64)
    dx = np.diff(k[..., 0], axis=1).mean(axis=1)
    dy = np.diff(k[..., 1], axis=1).mean(axis=1)
    lG = ridge(np.stack([dx, dy], -1).astype(np.complex128), lth, lamda)
    return {"x": sla.expm(lG[0].reshape(nc, nc)), "y": sla.expm(lG[1].reshape(nc, nc)), "z": None}


def logm_eig(G: np.ndarray) -> np.ndarray:
    w, V = np.linalg.eig(np.asarray(G, dtype=np.complex128))
    return V @ np.diag(np.log(w)) @ np.linalg.inv(V)


def powers(G: np.ndarray, exps: np.ndarray) -> np.ndarray:
    L = logm_eig(G)
    return np.stack([sla.expm(float(p) * L) for p in exps])


def deltas(nsteps: int, stepsize: float) -> np.ndarray:
    d = np.arange(nsteps).astype(np.float32)
    return ((d - np.float32((nsteps - 1) / 2)) * np.float32(stepsize)).astype(np.float32)


def table(calib, lamda=0.01, precision=1, quant=None, axis_order="reference"):
    """(K, C, C) complex128 table, nsteps, ndim."""
    _, stepsize, nsteps = quant_params(0.5, precision, quant)
    ops = unit_operators(calib, lamda)
    ex = deltas(nsteps, stepsize)
    Gx, Gy = powers(ops["x"], ex), powers(ops["y"], ex)
    if ops["z"] is None:
        t = np.einsum("jab,ibc->ijac", Gx, Gy).reshape(-1, *Gx.shape[-2:])
        return t, nsteps, 2
    Gz = powers(ops["z"], ex)
    # reference pairing: Gx (trained along the calibration's last axis) with
    # the axis-0 shift; "corrected": the axis-0 operator Gz with it
    first, last = (Gz, Gx) if axis_order == "corrected" else (Gx, Gz)
    # t[i, j, k] = first[k] @ Gy[j] @ last[i]
    FY = np.einsum("kab,jbc->jkac", first, Gy)
    t = np.einsum("jkac,icd->ijkad", FY, last).reshape(-1, *Gx.shape[-2:])
    return t, nsteps, 3
