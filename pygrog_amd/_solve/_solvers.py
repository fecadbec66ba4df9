"""CG and LSMR on duck-typed operators (reference src/pygrog/_solve/_solvers.py).

PyGROG convention: ``forward = A`` (image -> k-space), ``adjoint = A^H``.
CG solves ``(A^H A + damp I) x = A^H b``.  Its scalars (<r, z>, <p, Ap>,
alpha, beta, residual norms) are accumulated in float64 by default
(``scalar_dtype``), while the vectors and operator applications stay in the
data precision; ``scalar_dtype=torch.float32`` reproduces the reference's
complex64 ``vdot`` scalars (:117-134).
"""

from __future__ import annotations

__all__ = ["cg", "lsmr", "solve"]

import math

import torch


def _normalize_kspace_shape(op, b: torch.Tensor) -> torch.Tensor:
    """Reshape flat ``(..., n_samples)`` into ``(..., *natural_shape)`` so
    ``b - op.forward(x)`` aligns (reference :24-43)."""
    nat = getattr(op, "natural_shape", None)
    if nat is None or len(nat) <= 1:
        return b
    nat = tuple(int(s) for s in nat)
    n = 1
    for s in nat:
        n *= s
    if b.shape[-1] != n:
        return b
    return b.reshape(*b.shape[:-1], *nat)


def _vdot_c64_f64(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """<a, b> of two complex64 device vectors summed in float64 by the HIP
    reduction (no complex128 copies); a (2,) float64 device tensor."""
    from .. import _lib

    a = a.reshape(-1).resolve_conj().contiguous()
    b = b.reshape(-1).resolve_conj().contiguous()
    ws = torch.empty(int(_lib.load().pg_vdot_workspace_bytes()), dtype=torch.uint8, device=a.device)
    out = torch.empty(2, dtype=torch.float64, device=a.device)
    _lib.call("pg_vdot_c64_f64", _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(ws), _lib.ptr(out),
              _lib.stream(a.device))
    return out


def _dot(a: torch.Tensor, b: torch.Tensor, dt) -> torch.Tensor:
    """Re <a, b> accumulated in ``dt`` (a real dtype)."""
    ct = torch.complex128 if dt == torch.float64 else torch.complex64
    a = a.reshape(-1)
    b = b.reshape(-1)
    if (dt == torch.float64 and a.dtype == torch.complex64 and b.dtype == torch.complex64 and a.is_cuda
            and b.device == a.device):
        return _vdot_c64_f64(a, b)[0]
    if a.is_complex() or b.is_complex():
        return torch.vdot(a.to(ct), b.to(ct)).real
    return torch.dot(a.to(dt), b.to(dt))


def _norm(a: torch.Tensor, dt) -> torch.Tensor:
    if dt == torch.float64 and a.dtype == torch.complex64 and a.is_cuda:
        return torch.sqrt(_vdot_c64_f64(a, a)[0])
    ct = torch.complex128 if dt == torch.float64 else torch.complex64
    return torch.linalg.vector_norm(a.to(ct) if a.is_complex() else a.to(dt))


def cg(op, b, *, damp: float = 0.0, x0=None, max_iter: int = 20, tol: float = 1e-6, preconditioner=None,
       callback=None, scalar_dtype=torch.float64):
    """Conjugate gradients on the normal equations (reference :49-139)."""
    b = _normalize_kspace_shape(op, b)
    AHb = op.adjoint(b)
    AHA = getattr(op, "normal", None)
    if AHA is None:
        def AHA(v):
            return op.adjoint(op.forward(v))

    dt = scalar_dtype
    x = torch.zeros_like(AHb) if x0 is None else x0.clone()
    r = AHb - AHA(x) - damp * x
    z = preconditioner.apply(r) if preconditioner is not None else r
    p = z.clone()
    rz_old = _dot(r, z, dt)
    r0 = _norm(r, dt) + 1e-30
    # On a HIP device the residual test of iteration k is read only after
    # iteration k + 1 has been enqueued: its ratio goes to pinned host memory
    # asynchronously, so the host never waits for an idle device.  If
    # iteration k had converged, its saved iterate is returned and the one
    # enqueued after it is dropped: the same result and stopping rule as
    # testing every iteration synchronously.
    pipelined = x.is_cuda and isinstance(r0, torch.Tensor) and r0.is_cuda
    if pipelined:
        rel_host = torch.empty(2, dtype=torch.float64, pin_memory=True)
        pending = None  # (k, x_k, event, slot)

    def _settle(item):
        k_, x_, ev_, slot_ = item
        ev_.synchronize()
        rel_ = float(rel_host[slot_])
        if callback is not None:
            callback(k_, x_, rel_)
        return rel_ < tol

    if (pipelined and preconditioner is None and dt == torch.float64 and x.dtype == torch.complex64
            and p.dtype == torch.complex64 and x.numel() > 0):
        return _cg_fused(AHA, x, r, p, rz_old, r0, damp, max_iter, tol, callback, rel_host)

    for k in range(max_iter):
        Ap = AHA(p)
        if damp != 0.0:
            Ap = Ap + damp * p
        alpha = rz_old / (_dot(p, Ap, dt) + 1e-30)
        a = alpha.to(x.real.dtype if x.is_complex() else x.dtype)
        x = x + a * p
        r = r - a * Ap
        # ||r||^2 as <r, r>: the residual test, and <r, z> itself when z = r
        rr = _dot(r, r, dt)
        if pipelined:
            slot = k % 2
            rel_host[slot].copy_(torch.sqrt(rr) / r0, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            if pending is not None and _settle(pending):
                return pending[1]
            pending = (k, x, ev, slot)
        else:
            rel = float(torch.sqrt(rr) / r0)
            if callback is not None:
                callback(k, x, rel)
            if rel < tol:
                break
        z = preconditioner.apply(r) if preconditioner is not None else r
        rz_new = _dot(r, z, dt) if preconditioner is not None else rr
        beta = rz_new / (rz_old + 1e-30)
        p = z + beta.to(a.dtype) * p
        rz_old = rz_new
    if pipelined and pending is not None:
        _settle(pending)
        return pending[1]
    return x


def _cg_fused(AHA, x, r, p, rz_old, r0, damp, max_iter, tol, callback, rel_host):
    """The loop of ``cg`` for complex64 vectors on a HIP device with FP64
    scalars and no preconditioner: each iteration's vector and scalar work in
    two library calls (``pg_cg_step_xr``: <p, Ap>, alpha, x' = x + a p,
    r -= a Ap, <r, r>, the residual ratio and beta; ``pg_cg_step_p``:
    p = r + b p) with the scalars kept on the device.  The arithmetic is the
    generic loop's, rounding for rounding, so the iterates are its bit for
    bit; the residual test is read one iteration late as there.  Two x
    buffers alternate, so iterate k is intact when its test is read."""
    from .. import _lib

    dev = x.device
    st = _lib.stream(dev)
    shape = x.shape
    n = x.numel()
    xs = [x.reshape(-1).contiguous().clone(), torch.empty(n, dtype=x.dtype, device=dev)]
    r = r.reshape(-1).contiguous().clone()
    p = p.reshape(-1).contiguous().clone()
    sc = torch.zeros(8, dtype=torch.float64, device=dev)
    sc[0] = rz_old
    sc[3] = r0
    ws = torch.empty(int(_lib.load().pg_vdot_workspace_bytes()), dtype=torch.uint8, device=dev)
    pending = None  # (k, x_k, event, slot)

    def settle(item):
        k_, x_, ev_, slot_ = item
        ev_.synchronize()
        rel_ = float(rel_host[slot_])
        if callback is not None:
            callback(k_, x_.reshape(shape), rel_)
        return rel_ < tol

    cur = 0
    for k in range(max_iter):
        Ap = AHA(p.reshape(shape))
        if damp != 0.0:
            Ap = Ap + damp * p.reshape(shape)
        Ap = Ap.reshape(-1).contiguous()
        _lib.call("pg_cg_step_xr", _lib.ptr(p), _lib.ptr(Ap), _lib.ptr(xs[cur]), _lib.ptr(xs[1 - cur]), _lib.ptr(r),
                  n, _lib.ptr(sc), _lib.ptr(ws), st)
        cur = 1 - cur
        slot = k % 2
        rel_host[slot].copy_(sc[4], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        if pending is not None and settle(pending):
            return pending[1].reshape(shape)
        pending = (k, xs[cur], ev, slot)
        _lib.call("pg_cg_step_p", _lib.ptr(r), _lib.ptr(p), n, _lib.ptr(sc), st)
    if pending is not None:
        settle(pending)
        return pending[1].reshape(shape)
    return xs[cur].reshape(shape)


def lsmr(op, b, *, damp: float = 0.0, x0=None, max_iter: int = 20, atol: float = 1e-6, btol: float = 1e-6,
         callback=None):
    """LSMR for ``min ||A x - b||^2 + damp^2 ||x||^2`` (Fong & Saunders,
    SIAM J. Sci. Comput. 33(5), 2011, Algorithm 1; reference :145-262).

    Stops when ``|zetabar| / ||b|| < atol + btol`` (the reference's test).
    """
    b = _normalize_kspace_shape(op, b)
    if x0 is not None:
        b = b - op.forward(x0)
    beta = float(torch.linalg.vector_norm(b).item())
    if beta == 0.0:
        return x0.clone() if x0 is not None else op.adjoint(b)
    u = b / beta
    v = op.adjoint(u)
    alpha = float(torch.linalg.vector_norm(v).item())
    if alpha == 0.0:
        return x0.clone() if x0 is not None else torch.zeros_like(v)
    v = v / alpha

    zetabar = alpha * beta
    alphabar = alpha
    rho = 1.0
    rhobar = 1.0
    cbar = 1.0
    sbar = 0.0
    h = v.clone()
    hbar = torch.zeros_like(v)
    x = torch.zeros_like(v) if x0 is None else x0.clone()
    normb = beta
    for k in range(max_iter):
        # bidiagonalisation
        u = op.forward(v) - alpha * u
        beta = float(torch.linalg.vector_norm(u).item())
        if beta > 0:
            u = u / beta
            v = op.adjoint(u) - beta * v
            alpha = float(torch.linalg.vector_norm(v).item())
            if alpha > 0:
                v = v / alpha
        # rotation P_k (damping), then Q_k
        alphahat = math.hypot(alphabar, damp)
        rhoold = rho
        rho = math.hypot(alphahat, beta)
        c = alphahat / rho
        s = beta / rho
        thetanew = s * alpha
        alphabar = c * alpha
        # rotation Qbar_k
        rhobarold = rhobar
        thetabar = sbar * rho
        rhotemp = cbar * rho
        rhobar = math.hypot(rhotemp, thetanew)
        cbar = rhotemp / rhobar
        sbar = thetanew / rhobar
        zeta = cbar * zetabar
        zetabar = -sbar * zetabar
        # updates
        hbar = h - (thetabar * rho / (rhoold * rhobarold)) * hbar
        x = x + (zeta / (rho * rhobar)) * hbar
        h = v - (thetanew / rho) * h
        rel = abs(zetabar) / (normb + 1e-30)
        if callback is not None:
            callback(k, x, rel)
        if rel < atol + btol:
            break
    return x


def solve(op, b, *, method: str = "cg", **kwargs):
    """Dispatch to :func:`cg` or :func:`lsmr` (reference :268-302)."""
    m = method.lower()
    if m == "cg":
        return cg(op, b, **kwargs)
    if m == "lsmr":
        return lsmr(op, b, **kwargs)
    raise ValueError(f"Unknown solver method: {method!r}. Use 'cg' or 'lsmr'.")
