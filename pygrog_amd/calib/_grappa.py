"""GRAPPA operator table for GROG, built in FP64 on the GPU.

Follows the reference algorithm (src/pygrog/calib/_grappa.py):
  * calibration normalised by its Frobenius norm                 (:275-287)
  * unit-shift operators G = X^T from the Tikhonov least squares
    min ||S X - T||^2 + lamda ||X||^2, S/T = shifted ACR rows     (:16-59, 326-368)
  * fractional powers G^p = expm(p * logm(G)), logm = V diag(log w) V^-1,
    evaluated as V diag(w^p) V^-1 for all axes in one eigendecomposition
                                                                 (:62-72, 184-208)
  * table[i1*n + i0] = Gx[i0] @ Gy[i1]            (2D)
    table[i2*n^2 + i1*n + i0] = Gx[i0] @ Gy[i1] @ Gz[i2]  (3D)   (:131-181)
with the exponents the reference uses: p_i = (i - (n-1)/2) * stepsize in
float32 (:100-103).

Where the reference computes in complex64 (LAPACK gelsd on the augmented
system, complex128 eig cast back to complex64, complex64 expm and products),
everything here runs in complex128 on the GPU and is cast to complex64 only
at the end.  The least squares is solved by QR of the augmented system
``[S; sqrt(lamda) I] X = [T; 0]`` (same minimiser as gelsd).  If this torch
build has no GPU kernel for a dense factorisation (eig needs MAGMA on ROCm),
that one factorisation of the C x C (or M x C) matrix runs in FP64 on the host
and ``KernelTableResult.host_steps`` records it.

3D axis pairing (reference quirk): ``_grappa_op_3d`` trains ``Gx`` along the
calibration's last axis while the quantised shift ``d0`` is grid axis 0, so
the reference table pairs ``Gx`` (axis-2 operator) with ``d0``.  The default
``axis_order="reference"`` reproduces that; ``"corrected"`` builds
``G_ax0^{d0} @ G_ax1^{d1} @ G_ax2^{d2}`` instead.
"""

from __future__ import annotations

__all__ = ["KernelTable", "KernelTableResult", "build_table_finish", "build_table_launch", "compose_table",
           "grappa_power", "grappa_powers", "lstsq", "prepare_grappa_table", "train_grappa"]

import functools
import math
from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy as np
import torch

from .. import _lib


@dataclass
class KernelTableResult:
    table: torch.Tensor  # (K, C, C) complex64, on the compute device
    nsteps: int
    ndim: int
    stepsize: float
    host_steps: list = field(default_factory=list)
    table64: torch.Tensor | None = None  # complex128 table (kept on request)
    # the FP64 per-axis powers the table is assembled from (small: ndim x
    # nsteps x C x C), kept so the complex128 table can be rebuilt on demand
    powers64: tuple | None = None
    axis_order: str = "reference"


def _to_dev(x, device, dtype):
    t = torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x)
    return t.to(device=device, dtype=dtype)


# The Cartesian axis fits go through the normal equations (A^H A + lamda I)
# X = A^H b when lamda is at least this: the calibration is scaled to unit
# norm, so the system's condition number is at most about (1 + lamda) /
# lamda and FP64 keeps the operators to far better than the complex64 table
# needs (1e-15 relative to QR at lamda = 1e-3).  Below it, and for the
# radial self-calibration (whose per-spoke operators go through a matrix
# logarithm that amplifies any error), QR.
NORMAL_EQ_MIN_LAMDA = 1e-8


def lstsq(A, b, lamda: float = 0.0, host_steps: list | None = None, normal: bool = False):
    """Tikhonov least squares ``min |A X - b|^2 + lamda |X|^2`` (reference
    _grappa.py:16-59, which solves ``[A; sqrt(lamda) I] X = [b; 0]`` by QR).
    Batched over leading dims.  With ``normal`` and ``lamda >=
    NORMAL_EQ_MIN_LAMDA`` -- for a unit-norm A, whose system is then well
    conditioned -- the Gram products are formed on A's device (two GEMMs)
    and the C x C system is solved on the host: on a device the QR of a few
    thousand rows by C columns costs milliseconds of solver launches
    (measured on MI355X: 1.5 ms for a 16^3 x 16-coil region's three axes).
    Otherwise QR, as the reference."""
    A_t = torch.as_tensor(A)
    b_t = torch.as_tensor(b)
    N = A_t.shape[-1]
    if normal and lamda >= NORMAL_EQ_MIN_LAMDA:
        Ah = A_t.mH
        M = Ah @ A_t + lamda * torch.eye(N, dtype=A_t.dtype, device=A_t.device)
        R = Ah @ b_t
        return torch.linalg.solve(M.cpu(), R.cpu()).to(A_t.device)
    if lamda > 0:
        I_reg = math.sqrt(lamda) * torch.eye(N, dtype=A_t.dtype, device=A_t.device)
        if A_t.dim() > 2:
            I_reg = I_reg.expand(*A_t.shape[:-2], N, N)
        zeros = torch.zeros(*A_t.shape[:-2], N, b_t.shape[-1], dtype=b_t.dtype, device=b_t.device)
        A_aug = torch.cat([A_t, I_reg], dim=-2)
        b_aug = torch.cat([b_t, zeros], dim=-2)
    else:
        A_aug, b_aug = A_t, b_t
    try:
        Q, R = torch.linalg.qr(A_aug, mode="reduced")
        rhs = Q.mH @ b_aug
        return torch.linalg.solve_triangular(R, rhs, upper=True)
    except RuntimeError:
        if host_steps is not None:
            host_steps.append("lstsq")
        Q, R = torch.linalg.qr(A_aug.cpu(), mode="reduced")
        sol = torch.linalg.solve_triangular(R, Q.mH @ b_aug.cpu(), upper=True)
        return sol.to(A_t.device)


def _eig(A128: torch.Tensor, host_steps: list | None = None):
    try:
        return torch.linalg.eig(A128)
    except (RuntimeError, NotImplementedError):
        if host_steps is not None:
            host_steps.append("eig")
        vals, vecs = torch.linalg.eig(A128.cpu())
        return vals.to(A128.device), vecs.to(A128.device)


def _matrix_logm(A: torch.Tensor, host_steps: list | None = None) -> torch.Tensor:
    """logm(A) = V diag(log w) V^-1 in complex128 (reference _grappa.py:62-72,
    without the cast back to complex64)."""
    vals, vecs = _eig(A.to(torch.complex128), host_steps)
    return vecs @ torch.diag_embed(torch.log(vals)) @ torch.linalg.inv(vecs)


def _powers_host(G: np.ndarray, exps: np.ndarray) -> np.ndarray:
    """V diag(w^p) V^-1 for a stack of C x C complex128 operators and
    float64 exponents, on the host: (A, C, C), (n,) -> (A, n, C, C)."""
    vals, vecs = np.linalg.eig(G)
    vinv = np.linalg.inv(vecs)
    wp = np.exp(exps[None, :, None] * np.log(vals)[:, None, :])  # (A, n, C)
    return (vecs[:, None] * wp[:, :, None, :]) @ vinv[:, None]


def grappa_powers(G_units, exponents, host_steps: list | None = None,  # noqa: ARG001
                  device=None) -> torch.Tensor:
    """Fractional powers of a stack of unit operators, (A, C, C) -> (A, n, C, C)
    complex128 on ``device`` (default: the operators' device), one batched
    eigendecomposition for all of them.

    The reference forms G^p = expm(p logm(G)) with logm(G) = V diag(log w)
    V^-1 from G = V diag(w) V^-1 (_grappa.py:62-72, 184-208); that
    exponential is V diag(exp(p log w)) V^-1, evaluated here directly (same
    principal logarithm of the eigenvalues, no matrix exponential).  The
    operators are C x C (C coils), so the whole evaluation runs on the host
    (LAPACK through numpy) and only the powers go to the device: the
    device eigensolver spends milliseconds in launches and host round
    trips (measured on MI355X: 3 x 16 x 16 with 11 exponents, 3.4 ms on the
    device, 0.3 ms on the host; same values to 1e-15).  The upload goes
    through page-locked memory without waiting for the device's stream."""
    if isinstance(G_units, torch.Tensor):
        home = G_units.device
        Gh = G_units.detach().to(torch.complex128).cpu().numpy()
    else:
        home = torch.device("cpu")
        Gh = np.asarray(G_units, dtype=np.complex128)
    # exponents in float32, as the reference computes them
    if isinstance(exponents, torch.Tensor):
        exps = exponents.detach().to(torch.float32).cpu().numpy().astype(np.float64)
    else:
        exps = np.asarray(exponents, dtype=np.float32).astype(np.float64)
    P = torch.from_numpy(_powers_host(Gh, exps))
    dev = torch.device(device) if device is not None else home
    if dev.type == "cpu":
        return P
    return P.pin_memory().to(dev, non_blocking=True)


def grappa_power(G_unit: torch.Tensor, exponents, host_steps: list | None = None) -> torch.Tensor:
    """``G^p = expm(p logm(G))`` of one unit operator for every exponent,
    (n, C, C) complex128 (reference :184-208)."""
    return grappa_powers(torch.as_tensor(G_unit).unsqueeze(0), exponents, host_steps)[0]


def prepare_grappa_table(Gx, Gy, Gz, nsteps: int, ndim: int, axis_order: str = "reference") -> torch.Tensor:  # noqa: ARG001
    """Outer-product operator table (reference _grappa.py:131-181)."""
    Gx = torch.as_tensor(Gx)
    Gy = torch.as_tensor(Gy)
    if ndim == 2:
        # table[i * nsteps + j] = Gx[j] @ Gy[i]
        return (Gx[None] @ Gy[:, None]).reshape(-1, *Gx.shape[-2:])
    if ndim == 3:
        if Gz is None:
            raise ValueError("3D interpolation requires Z operator")
        Gz = torch.as_tensor(Gz)
        if axis_order == "corrected":
            # shift along grid axis 0 uses the axis-0 operator (Gz here)
            first, last = Gz, Gx
        else:
            first, last = Gx, Gz
        # table[i * nsteps^2 + j * nsteps + k] = first[k] @ Gy[j] @ last[i]
        return (first[None, None] @ Gy[None, :, None] @ last[:, None, None]).reshape(-1, *Gx.shape[-2:])
    raise ValueError(f"GROG interpolation only supports 2D or 3D data, got {ndim}D")


def _grappa_op_2d(calib, lamda, host_steps):
    calib = calib.movedim(0, -1)
    nc = calib.shape[-1]
    Sy = calib[:, :-1, :].reshape(-1, nc)
    Sx = calib[:-1, ...].reshape(-1, nc)
    Ty = calib[:, 1:, :].reshape(-1, nc)
    Tx = calib[1:, ...].reshape(-1, nc)
    Gy = lstsq(Sy, Ty, lamda, host_steps, normal=True).mT
    Gx = lstsq(Sx, Tx, lamda, host_steps, normal=True).mT
    return Gy, Gx


def _grappa_op_3d(calib, lamda, host_steps):
    calib = calib.movedim(0, -1)
    nc = calib.shape[-1]
    Sz = calib[:-1, :, :, :].reshape(-1, nc)
    Sy = calib[:, :-1, :, :].reshape(-1, nc)
    Sx = calib[:, :, :-1, :].reshape(-1, nc)
    Tz = calib[1:, :, :, :].reshape(-1, nc)
    Ty = calib[:, 1:, :, :].reshape(-1, nc)
    Tx = calib[:, :, 1:, :].reshape(-1, nc)
    if Sz.shape == Sy.shape == Sx.shape:
        # a cubic region: the three systems in one batched solve
        G = lstsq(torch.stack([Sz, Sy, Sx]), torch.stack([Tz, Ty, Tx]), lamda, host_steps, normal=True).mT
        return G[0], G[1], G[2]
    Gz = lstsq(Sz, Tz, lamda, host_steps, normal=True).mT
    Gy = lstsq(Sy, Ty, lamda, host_steps, normal=True).mT
    Gx = lstsq(Sx, Tx, lamda, host_steps, normal=True).mT
    return Gz, Gy, Gx


def _radial_grappa_op(calib, lamda, coords, host_steps):
    """Unit operators from radial data (reference _grappa.py:290-323), FP64.

    One operator per spoke maps each readout sample to the next (least
    squares over the spoke), its matrix log is fitted as dx * log Gx +
    dy * log Gy over all spokes (dx, dy = the spoke's mean readout step),
    and the exponentials are returned as (Gy, Gx).  The logs are flattened
    column-major and unflattened row-major, as in the reference, which turns
    the row-vector fit into the column-vector operators the table uses."""
    calib = calib.movedim(0, -1)  # (nr, ns, nc)
    nr, _ns, nc = calib.shape
    c = coords.to(calib.device, torch.float64)
    xcoord = c[..., 0].T
    ycoord = c[..., 1].T
    S = calib[:, :-1, ...]
    T = calib[:, 1:, ...]
    Gtheta = lstsq(S, T, lamda, host_steps)  # (nr, nc, nc)
    lGtheta = _matrix_logm(Gtheta, host_steps).permute(0, 2, 1).reshape(nr, nc * nc).contiguous()
    dx = torch.diff(xcoord, dim=0).mean(dim=0)
    dy = torch.diff(ycoord, dim=0).mean(dim=0)
    dxy = torch.stack([dx, dy], dim=-1).to(lGtheta.dtype)
    lG = lstsq(dxy, lGtheta, lamda, host_steps)  # (2, nc^2)
    lGx = lG[0].reshape(nc, nc)
    lGy = lG[1].reshape(nc, nc)
    return torch.linalg.matrix_exp(lGy), torch.linalg.matrix_exp(lGx)


def _gram_fit(calib: torch.Tensor, lamda: float, on_host: bool = False) -> torch.Tensor:
    """Unit operators of a Cartesian calibration ``(C, *dims)`` complex128
    on a device, one per axis (axis 0 first), by the normal equations: the
    FP64 products G_d = A_d^H A_d, X_d = A_d^H T_d and |x|^2 from one kernel
    (``pg_grappa_gram_f64``), then ((G_d + lamda |x|^2 I)^-1 X_d)^T solved on
    the host -- the fit of the calibration scaled to unit norm, as the
    reference does (_grappa.py:211-287), without forming the source and
    target matrices.  ``on_host``: leave the operators on the host (where the
    table build takes their eigendecompositions next)."""
    ops = _gram_ops(_gram_launch(calib, lamda))
    return ops if on_host else ops.to(calib.device)


def _gram_launch(calib: torch.Tensor, lamda: float) -> SimpleNamespace:
    """Enqueue the Gram products of ``_gram_fit`` and their copy into
    page-locked host memory on the current stream; ``_gram_ops`` waits for
    them and solves."""
    C = int(calib.shape[0])
    dims = [int(n) for n in calib.shape[1:]]
    nd = len(dims)
    x = calib.contiguous()
    out = torch.empty(2 * nd * C * C + C, dtype=torch.complex128, device=x.device)
    _lib.call("pg_grappa_gram_f64", _lib.ptr(x), C, nd, _lib.dims(dims), _lib.ptr(out), _lib.stream(x.device))
    host = torch.empty(out.shape, dtype=out.dtype, pin_memory=True)
    host.copy_(out, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    return SimpleNamespace(host=host, event=ev, C=C, nd=nd, lamda=lamda, device=x.device)


def _gram_ops(g: SimpleNamespace) -> torch.Tensor:
    """(nd, C, C) complex128 unit operators on the host from ``_gram_launch``."""
    g.event.synchronize()
    C, nd = g.C, g.nd
    h = g.host.numpy()
    n2 = float(h[2 * nd * C * C:].real.sum())
    M = h[:2 * nd * C * C].reshape(nd, 2, C, C)
    return torch.from_numpy(np.linalg.solve(M[:, 0] + (g.lamda * n2) * np.eye(C), M[:, 1]).transpose(0, 2, 1))


def _gram_applies(calib: torch.Tensor, dev: torch.device, coords, lamda: float) -> bool:
    """Whether ``train_grappa`` fits ``calib`` by the normal equations."""
    return (dev.type != "cpu" and coords is None and lamda >= NORMAL_EQ_MIN_LAMDA and calib.ndim in (3, 4)
            and calib.shape[0] <= 256 and min(calib.shape[1:]) >= 2)


def _units_of(ops: torch.Tensor, ndim: int) -> SimpleNamespace:
    """Axis operators (axis 0 first) as the fit's x / y / z units."""
    if ndim == 2:
        return SimpleNamespace(x=ops[0], y=ops[1], z=None)
    return SimpleNamespace(x=ops[2], y=ops[1], z=ops[0])


def train_grappa(train_data, lamda: float | None = None, coords=None, *, device=None, host_steps: list | None = None,
                 host_units: bool = False) -> SimpleNamespace:
    """Unit-shift GROG operators (reference _grappa.py:211-287), FP64.

    Cartesian calibration ``(coils, *2 or 3 spatial dims)``, or with
    ``coords`` (spokes, readout, 2) radial self-calibration from
    ``(coils, spokes, readout)`` data (lamda defaults to 0 then, as in the
    reference).  ``host_units``: the normal-equation fit may leave the
    operators on the host."""
    if lamda is None:
        lamda = 0.01 if coords is None else 0.0
    dev = torch.device(device) if device is not None else torch.device("cpu")
    calib = _to_dev(train_data, dev, torch.complex128)
    if _gram_applies(calib, dev, coords, lamda):
        return _units_of(_gram_fit(calib, lamda, on_host=host_units), calib.ndim - 1)
    calib = calib / torch.linalg.norm(calib)
    if coords is not None:
        c = torch.as_tensor(np.asarray(coords) if not isinstance(coords, torch.Tensor) else coords)
        if c.shape[-1] != 2 or calib.ndim != 3 or tuple(c.shape[:-1]) != tuple(calib.shape[1:]):
            raise ValueError(f"radial self-calibration takes (coils, spokes, readout) data and (spokes, readout, 2) "
                             f"coords, got {tuple(calib.shape)} and {tuple(c.shape)}")
        gy, gx = _radial_grappa_op(calib, lamda, c, host_steps)
        return SimpleNamespace(x=gx, y=gy, z=None)
    ndim = calib.ndim - 1
    if ndim == 2:
        gy, gx = _grappa_op_2d(calib, lamda, host_steps)
        return SimpleNamespace(x=gx, y=gy, z=None)
    if ndim == 3:
        gz, gy, gx = _grappa_op_3d(calib, lamda, host_steps)
        return SimpleNamespace(x=gx, y=gy, z=gz)
    raise ValueError(f"calibration must be (coils, *2 or 3 spatial dims), got {tuple(calib.shape)}")


def quant_steps(radius: float, precision: int | None = None, quant: int | None = None):
    """(pfac, stepsize, nsteps) of the shift quantisation.

    ``precision`` p (reference, _grog.py:406-408 and _grappa.py:100-101):
    pfac = 10**p, stepsize = 10**-p, nsteps = int(2 radius / stepsize + 1)
    in Python double arithmetic.  ``quant`` q (extension for 1/q-grid
    steps): pfac = q, stepsize = 1/q, nsteps = q + 1.
    """
    if quant is not None:
        q = int(quant)
        if q < 1:
            raise ValueError("quant must be a positive integer")
        return float(q), 1.0 / q, q + 1
    p = int(precision if precision is not None else 1)
    stepsize = 10 ** (-p)
    return 10.0**p, stepsize, int(2 * radius / stepsize + 1)


def build_table(train_data, radius: float, precision: int | None = 1, lamda: float = 0.01, *, quant=None,
                device=None, axis_order: str = "reference", keep_fp64: bool = False) -> KernelTableResult:
    """Operator table on ``device`` (complex64 result of an FP64 build)."""
    host_steps: list = []
    # the unit operators stay on the host when the fit leaves them there: their
    # eigendecompositions run there, and only the powers go to the device
    kern = train_grappa(train_data, lamda, device=device, host_steps=host_steps, host_units=True)
    return _table_from_units(kern, radius, precision, quant, device, axis_order, keep_fp64, host_steps)


def build_table_launch(train_data, radius: float, precision: int | None = 1, lamda: float = 0.01, *, quant=None,
                       device=None, axis_order: str = "reference", keep_fp64: bool = False) -> SimpleNamespace:
    """``build_table`` in two steps: this one enqueues the device part of a
    Cartesian normal-equation fit (calibration cast and Gram products, then
    their copy to the host) on the current stream and returns;
    ``build_table_finish`` waits for them and does the rest, so the caller
    can enqueue other work (a plan compiler) in between.  Other fits run
    here in full."""
    dev = torch.device(device) if device is not None else torch.device("cpu")
    args = SimpleNamespace(radius=radius, precision=precision, quant=quant, device=device, axis_order=axis_order,
                           keep_fp64=keep_fp64, gram=None, res=None, ndim=None)
    calib = _to_dev(train_data, dev, torch.complex128)
    if _gram_applies(calib, dev, None, lamda):
        args.gram = _gram_launch(calib, lamda)
        args.ndim = calib.ndim - 1
    else:
        args.res = build_table(train_data, radius, precision, lamda, quant=quant, device=device,
                               axis_order=axis_order, keep_fp64=keep_fp64)
    return args


def build_table_finish(p: SimpleNamespace) -> KernelTableResult:
    """The table of a ``build_table_launch``."""
    if p.res is not None:
        return p.res
    kern = _units_of(_gram_ops(p.gram), p.ndim)
    return _table_from_units(kern, p.radius, p.precision, p.quant, p.device, p.axis_order, p.keep_fp64, [])


@functools.lru_cache(maxsize=64)
def _exponents(nsteps: int, stepsize: float) -> np.ndarray:
    """The power exponents (i - (n-1)/2) * stepsize, computed in float32 as
    the reference does (_grappa.py:100-103), as float64 (read-only)."""
    deltas = torch.arange(nsteps).float()
    deltas = (deltas - (nsteps - 1) / 2) * stepsize
    out = deltas.numpy().astype(np.float64)
    out.flags.writeable = False
    return out


def _table_from_units(kern, radius, precision, quant, device, axis_order, keep_fp64, host_steps) -> KernelTableResult:
    pfac, stepsize, nsteps = quant_steps(radius, precision, quant)
    deltas = _exponents(nsteps, stepsize)
    units = [kern.x, kern.y] + ([kern.z] if kern.z is not None else [])
    dev = torch.device(device) if device is not None else units[0].device
    if all(isinstance(u, torch.Tensor) and u.device.type == "cpu" for u in units):
        stacked = np.stack([u.numpy() for u in units])
    else:
        stacked = torch.stack([torch.as_tensor(u) for u in units])
    powers = grappa_powers(stacked, deltas, host_steps, device=dev)
    Gx, Gy = powers[0], powers[1]
    if kern.z is not None:
        Gz = powers[2]
        ndim = 3
    else:
        Gz = None
        ndim = 2
    t32, t64 = compose_table(powers, ndim, nsteps, axis_order, keep_fp64)
    return KernelTableResult(t32, nsteps, ndim, stepsize, host_steps, t64,
                             powers64=(Gx, Gy, Gz), axis_order=axis_order)


def compose_table(powers: torch.Tensor, ndim: int, nsteps: int, axis_order: str = "reference",
                  keep_fp64: bool = False):
    """(complex64 table, complex128 table or None) from the stacked per-axis
    powers (ndim, nsteps, C, C) complex128: one library call
    (``pg_table_compose_f64``) on a HIP device for up to 32 coils, else the
    batched products of ``prepare_grappa_table``."""
    C = int(powers.shape[-1])
    if powers.device.type == "cuda" and C <= 32 and ndim in (2, 3):
        P = powers.contiguous()
        K = nsteps ** ndim
        t32 = torch.empty(K, C, C, dtype=torch.complex64, device=P.device)
        t64 = torch.empty(K, C, C, dtype=torch.complex128, device=P.device) if keep_fp64 else None
        _lib.call("pg_table_compose_f64", _lib.ptr(P), ndim, nsteps, C, 1 if axis_order == "corrected" else 0,
                  _lib.ptr(t32), _lib.ptr(t64), _lib.stream(P.device))
        return t32, t64
    Gz = powers[2] if ndim == 3 else None
    t64 = prepare_grappa_table(powers[0], powers[1], Gz, nsteps, ndim, axis_order=axis_order)
    return t64.to(torch.complex64).contiguous(), (t64 if keep_fp64 else None)


def KernelTable(train_data, radius: float, precision: int = 1, lamda: float = 0.01, *, quant=None, device=None,
                axis_order: str = "reference"):
    """Reference-compatible entry (``_grappa.py:75-128``): returns
    ``(table (K, C, C) complex64 numpy array, nsteps, ndim)``."""
    res = build_table(train_data, radius, precision, lamda, quant=quant, device=device, axis_order=axis_order)
    return res.table.cpu().numpy(), res.nsteps, res.ndim
