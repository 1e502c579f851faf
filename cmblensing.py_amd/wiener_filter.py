"""Lensed Wiener filter by preconditioned conjugate gradient on the GPU
(wiener_filter.py:10-105 of the reference).

Precision: the operator A (two LenseFlows and the FFT/rotation chain,
``cmbl_gradf_logpdf``) runs in the field's dtype (float32 for the mixed
configuration); the CG vectors x, r, z, p, the inner products, alpha and beta
are float64 (``cmbl_dot_f64``, ``cmbl_cg_update_xrz``, ``cmbl_cg_update_p``),
and the residual recomputation b - A x applies the operator in float64 (mixed
precision iterative refinement: the float32 operator's error, ~1e-6 relative,
no longer bounds the reachable residual).  Several right-hand sides (a
leading batch axis) are solved together, each with its own alpha, beta and
stopping test.

Stopping: the reference stops when dot(r, P^-1 r) < tol (absolute, tol = 0.1)
or after maxiter = 500 iterations (wiener_filter.py:21-22, 75-77); that is the
default here.  ``rtol`` adds a relative test sqrt(dot(r, r) / dot(b, b)) <
rtol that is only passed by a recomputed residual b - A x (so the reported
residual is the true one, to the operator's precision), and
``recompute_every`` also replaces the recursive residual every k iterations.
``refine_drop`` (with ``rtol``) replaces a right-hand side's recursive
residual whenever its relative residual has fallen by that factor since its
last replacement (a reliable-update rule): with a float32 operator the
recursion drifts from b - A x as the residual shrinks below the operator's
error, and without a replacement within ~5-6 decades it can break down
(alpha from a noise-dominated p.Ap).  The default for a float32 operator
with ``rtol`` is 1e-4, or 1e-5 beside a ``recompute_every`` cadence, so every
right-hand side stays within that reach whatever its convergence rate.
"""
from __future__ import annotations

from dataclasses import dataclass, field as dc_field
from typing import List, Optional

import torch

from . import _lib
from .constants import DEFAULT_NSTEPS
from .fields import Field
from .gradients import (WfOperators, _ncomp, gradf_apply, phi_handles, stack_matrices,
                        unstack_matrices)
from .matrix_operators import BlockTEB, DiagonalEB, DiagonalScalar, PixelMask, pinv


# Apply the operator and the recomputed residuals to the right-hand sides still
# active only (False: to every right-hand side, whose inactive ones then get
# alpha = beta = 0; the iterates are the same either way).
COMPACT_ACTIVE = True

# Default reliable-update factors for a float32 operator with a relative stop:
# a right-hand side's recursive residual is replaced by b - A x (float64) once
# its relative residual has fallen by this factor since the last replacement.
# With a fixed cadence (recompute_every) as well, the rule is a safety net for
# right-hand sides that converge faster than the cadence expects: 1e-5 (the
# recursion broke down between 5 and 6 decades past a replacement on the
# 512^2 Wiener filter; a cadence of 20 moves about 4).
REFINE_DROP_FP32 = 1e-4
REFINE_DROP_FP32_CADENCE = 1e-5


@dataclass
class CGInfo:
    iterations: List[int]
    converged: List[bool]
    res_history: List[List[float]] = dc_field(default_factory=list)   # dot(r, z) per iteration
    rel_history: List[List[float]] = dc_field(default_factory=list)   # ||r|| / ||b||
    final_true_rel_residual: Optional[List[float]] = None


class _Problem:
    """Geometry, plan, prepared phi and operator arrays of one solve."""

    def __init__(self, field: Field, phi, field_covariance, noise_covariance, mask, beam, n: int,
                 batch: int):
        self.C = _ncomp(field)
        m0 = getattr(field, field._matrix_names()[0])
        self.N = int(m0.shape[-2])
        self.B = batch
        self.n = n
        self.device = m0.device
        self.cdt = m0.dtype if m0.is_complex() else (
            torch.complex64 if m0.dtype == torch.float32 else torch.complex128)
        self.rdt = torch.float32 if self.cdt == torch.complex64 else torch.float64
        self.plan = _lib.get_plan(self.N, field.theta_pix, self.rdt, self.B * self.C, self.device)
        self.ops = WfOperators(field_covariance, noise_covariance, mask, beam, self.rdt,
                               self.device)
        self.phis = phi_handles(self.plan, phi, self.B, n)
        self._args = (field.theta_pix, phi, field_covariance, noise_covariance, mask, beam)
        self.plan64 = None
        self.mixed = self.rdt == torch.float32

    def A64(self, X64: torch.Tensor, idx: Optional[List[int]] = None) -> torch.Tensor:
        """A x with the float64 operator (residual recomputation of the mixed
        mode): a second plan, operator arrays and phi handles in float64.
        With idx, X64 holds only those right-hand sides (len(idx) of them)."""
        if self.plan64 is None:
            th, phi, Cf, Cn, M, B = self._args
            self.plan64 = _lib.get_plan(self.N, th, torch.float64, self.B * self.C, self.device)
            self.ops64 = WfOperators(Cf, Cn, M, B, torch.float64, self.device)
            self.phis64 = phi_handles(self.plan64, phi, self.B, self.n)
        phis = self.phis64 if idx is None else [self.phis64[i] for i in idx]
        return gradf_apply(self.plan64, phis, X64, None, self.ops64, len(phis), self.C, self.n)

    def A(self, X_op: torch.Tensor, idx: Optional[List[int]] = None) -> torch.Tensor:
        """A x = gradf(x, d = 0) (wiener_filter.py:45-49), operator precision.
        With idx, X_op holds only those right-hand sides (len(idx) of them)."""
        phis = self.phis if idx is None else [self.phis[i] for i in idx]
        return gradf_apply(self.plan, phis, X_op, None, self.ops, len(phis), self.C, self.n)

    def A_active(self, X_op: torch.Tensor, idx: Optional[List[int]]) -> torch.Tensor:
        """A x for the right-hand sides in idx (all when idx is None), zero for
        the others: a request's maps go through the same kernels with the same
        arithmetic whatever batch they are in, so each applied one is the same
        bit for bit as in the full batch."""
        if idx is None:
            return self.A(X_op)
        sub = torch.tensor(idx, device=self.device)
        out = torch.zeros_like(X_op)
        out.index_copy_(0, sub, self.A(X_op.index_select(0, sub), idx))
        return out

    def dot(self, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        out = torch.empty(self.B, dtype=torch.float64, device=self.device)
        _lib.dot_f64(self.plan, torch.view_as_real(a), torch.view_as_real(b), self.B, self.C, out)
        return out


def get_b_vector(phi, data, field_covariance, noise_covariance, mask, beam,
                 n: int = DEFAULT_NSTEPS, like: Optional[Field] = None) -> Field:
    """b = -gradf_logpdf(0, d) (wiener_filter.py:28-33).  The forward LenseFlow
    of the zero field is skipped (it is exactly zero)."""
    tmpl = like if like is not None else data
    X, lead = stack_matrices(tmpl)
    prob = _Problem(tmpl, phi, field_covariance, noise_covariance, mask, beam, n, X.shape[0])
    D, _ = stack_matrices(data)
    D = D.to(prob.cdt if not prob.ops.pixel else prob.rdt)
    if D.shape[0] == 1 and prob.B > 1:
        D = D.expand(prob.B, *D.shape[1:]).contiguous()
    out = gradf_apply(prob.plan, prob.phis, None, D, prob.ops, prob.B, prob.C, n)
    return unstack_matrices(tmpl, -out, lead) if like is not None else \
        _as_fourier_like(data, -out, lead)


def _as_fourier_like(data: Field, X: torch.Tensor, lead) -> Field:
    from .constants import Basis
    f = unstack_matrices(data, X, lead)
    return f.replace(basis=Basis.FOURIER)


def hessian_logpdf_preconditioner(field_covariance, beam, mask, noise_covariance):
    """P = Cf^-1 + B M Cn^-1 M B (wiener_filter.py:35-43).  For a PixelMask the
    mask is replaced by its mean square (a Fourier-diagonal approximation)."""
    if isinstance(mask, PixelMask):
        return pinv(field_covariance) + beam * beam * pinv(noise_covariance) * mask.mean_square()
    return pinv(field_covariance) + beam * mask * pinv(noise_covariance) * mask * beam


def A_matrix_operator(field, phi, data, field_covariance, noise_covariance, mask, beam,
                      n: int = DEFAULT_NSTEPS) -> Field:
    """gradf_logpdf at d = 0 (wiener_filter.py:45-49)."""
    X, lead = stack_matrices(field)
    prob = _Problem(field, phi, field_covariance, noise_covariance, mask, beam, n, X.shape[0])
    return unstack_matrices(field, prob.A(X.to(prob.cdt)), lead)


def _cg_step(prob, x, r, z, p, Ap, pinv_arr, p_op, res_d, act_d, bb_d, rel_d, B):
    """One PCG iteration after the operator on the device (cmbl_cg_step):
    res_d is updated in place."""
    V = torch.view_as_real
    _lib.cg_step(prob.plan, V(x), V(r), V(z), V(p), V(Ap), pinv_arr, V(p_op), res_d, act_d, bb_d,
                 rel_d, B, prob.C)


def _host(t: torch.Tensor):
    """Wait point of a solve process: copies t into pinned host memory behind
    the work that produces it, yields an event recorded after the copy, then
    returns the host copy as float64 (one wait, no second device round
    trip)."""
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(t.device))
    yield ev
    return h.double()


def run_process(gen):
    """Run a solve process to completion on the calling thread."""
    try:
        ev = next(gen)
        while True:
            ev.synchronize()
            ev = gen.send(None)
    except StopIteration as stop:
        return stop.value


def conjugate_gradient_process(x0, data, phi, mask, beam, noise_covariance, field_covariance,
                               b_vector, preconditioner, maxiter: int = 500, tol: float = 1e-1, *,
                               n: int = DEFAULT_NSTEPS, rtol: Optional[float] = None,
                               recompute_every: Optional[int] = None, return_info: bool = False,
                               refine_drop="auto"):
    """PCG of wiener_filter.py:51-105, batched over right-hand sides, float64
    vectors and scalars around a float32 (or float64) operator, written as a
    process: a generator that enqueues device work and yields a HIP event at
    every point where the host needs a scalar (alpha, beta, the stop test), so
    a scheduler can interleave it with other work; ``run_process`` runs it to
    completion.  Returns what ``conjugate_gradient`` returns."""
    X0, lead = stack_matrices(x0)
    B = X0.shape[0]
    prob = _Problem(x0, phi, field_covariance, noise_covariance, mask, beam, n, B)
    dev = prob.device
    if refine_drop == "auto":
        refine_drop = (None if (rtol is None or not prob.mixed) else
                       REFINE_DROP_FP32_CADENCE if recompute_every else REFINE_DROP_FP32)
    x = X0.to(torch.complex128).contiguous().clone()
    bvec, _ = stack_matrices(b_vector)
    bvec = bvec.to(torch.complex128)
    if bvec.shape[0] == 1 and B > 1:
        bvec = bvec.expand(B, *bvec.shape[1:])
    bvec = bvec.contiguous()
    pinv_P = pinv(preconditioner)
    if isinstance(pinv_P, (DiagonalScalar, DiagonalEB, BlockTEB)):
        pinv_arr = pinv_P.per_mode().to(device=dev, dtype=torch.float64).contiguous()
    else:
        raise NotImplementedError("preconditioner must be DiagonalScalar, DiagonalEB or BlockTEB")

    r = torch.empty_like(x)
    z = torch.empty_like(x)
    ones = torch.ones(B, dtype=torch.float64, device=dev)
    true_r = [False] * B        # r[i] is a recomputed b - A x at the current x[i]

    def residual(xv, idx: Optional[List[int]] = None, x_zero: bool = False):
        """r = b - A x, z = P^-1 r for the right-hand sides in idx (all when
        None; the others keep r and z), all in float64 (with the float64
        operator in the mixed mode: iterative refinement).  Returns dot(r, z)
        per right-hand side (zero outside idx).  x_zero: x is all zeros, so
        A x = 0 is not applied."""
        sub = None if idx is None else torch.tensor(idx, device=dev)
        xs = xv if sub is None else xv.index_select(0, sub)
        k_ = xs.shape[0]
        if x_zero:
            Ax, plan_r = torch.zeros(xs.shape, dtype=prob.cdt, device=dev), prob.plan
        elif prob.mixed:
            Ax, plan_r = prob.A64(xs, idx), prob.plan64
        else:
            Ax, plan_r = prob.A(xs.to(prob.cdt), idx), prob.plan
        rs = (bvec if sub is None else bvec.index_select(0, sub)).clone()
        zs = torch.empty_like(rs)
        rzs = torch.empty(k_, dtype=torch.float64, device=dev)
        _lib.cg_update_xrz(plan_r, torch.view_as_real(torch.zeros_like(rs)),
                           torch.view_as_real(rs), torch.view_as_real(zs),
                           torch.view_as_real(xs.contiguous()), torch.view_as_real(Ax), pinv_arr,
                           ones[:k_], k_, prob.C, rzs)
        if sub is None:
            r.copy_(rs)
            z.copy_(zs)
            for i in range(B):
                true_r[i] = True
            return rzs
        r.index_copy_(0, sub, rs)
        z.index_copy_(0, sub, zs)
        for i in idx:
            true_r[i] = True
        return torch.zeros(B, dtype=torch.float64, device=dev).index_copy_(0, sub, rzs)

    # A x0 is skipped for x0 = 0 (every term of A is linear: exactly zero)
    x_nonzero = yield from _host(torch.any(x))
    res = residual(x, x_zero=not bool(x_nonzero))             # :61-69
    p = z.clone()
    p_op = p.to(prob.cdt).contiguous()
    bb = yield from _host(prob.dot(bvec, bvec))
    res_h = yield from _host(res)
    res_curr = res_h.clone()
    k = torch.zeros(B, dtype=torch.long)

    def rel_now():
        rr = yield from _host(prob.dot(r, r))
        return torch.sqrt(rr / torch.where(bb > 0, bb, torch.ones_like(bb)))

    rel = (yield from rel_now()) if rtol is not None else None

    # right-hand sides whose recursive residual has passed the relative test
    # and that wait, paused (alpha = beta = 0, no iteration counted), for the
    # FP64 recomputation that confirms or reactivates them
    verify = torch.zeros(B, dtype=torch.bool)

    def active_mask():
        act = (res_curr >= tol) & (k < maxiter)               # :75-77
        if rtol is not None:
            act &= (rel >= rtol) & ~verify
        return act

    info = CGInfo(iterations=[0] * B, converged=[False] * B,
                  res_history=[[float(v)] for v in res_h],
                  rel_history=[[float(v)] for v in rel] if rel is not None else [])
    act = active_mask()
    it = 0
    if rtol is None and not recompute_every and bool(act.any()):
        # The reference's stop rule only (absolute tol on dot(r, z), maxiter):
        # alpha, beta and the per-RHS active mask stay on the device (the same
        # float64 divisions, so the same iterates), and the host reads whether
        # any RHS is still active one iteration late, so the next iteration is
        # always queued while the device runs the current one.  The iteration
        # queued after the last active one has alpha = beta = 0 for every RHS:
        # it leaves x, the residuals and the counts unchanged.
        res_d = res_h.to(dev)
        cur_d = res_curr.to(dev)
        k_d = torch.zeros(B, dtype=torch.long, device=dev)
        act_d = (cur_d >= tol) & (k_d < maxiter)
        hist = []
        pending = None
        while True:
            Ap = prob.A(p_op)                                 # :82
            # :84-95 in one call: alpha, x, r, z, dot(r, z), beta, p, and
            # res = dot(r, z) for the active right-hand sides
            _cg_step(prob, x, r, z, p, Ap, pinv_arr, p_op, res_d, act_d, None, None, B)
            cur_d = res_d.clone()
            k_d = k_d + act_d.long()
            hist.append((act_d, cur_d))
            act_d = (cur_d >= tol) & (k_d < maxiter)          # :75-77
            flag = act_d.any()
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream(dev))
            if pending is not None:
                yield pending[0]
                if not bool(pending[1].cpu()):
                    break
            pending = (ev, flag)
        res_curr = cur_d.cpu()
        k = k_d.cpu()
        true_r[:] = [False] * B
        if hist:
            acts = torch.stack([h[0] for h in hist]).cpu()
            curs = torch.stack([h[1] for h in hist]).cpu()
            for a_row, c_row in zip(acts, curs):
                for i in range(B):
                    if bool(a_row[i]):
                        info.res_history[i].append(float(c_row[i]))
        act = torch.zeros(B, dtype=torch.bool)
    # Relative test and/or periodic refinement: alpha and beta stay on the
    # device (the same float64 divisions as on the host, so the same
    # iterates); the host reads dot(r, z) and the relative residual once per
    # iteration.  The operator and the recomputed residuals are applied to the
    # right-hand sides still active only.
    bb_d = torch.where(bb > 0, bb, torch.ones_like(bb)).to(dev)

    def rel_dev():
        return torch.sqrt(prob.dot(r, r) / bb_d)

    def record(a_row, cur_row, rel_row):
        for i in range(B):
            if bool(a_row[i]):
                info.res_history[i].append(float(cur_row[i]))
                if rel_row is not None:
                    info.rel_history[i].append(float(rel_row[i]))

    # the host's bookkeeping of an iteration (histories) is done while the
    # device runs the next one, after that one has been enqueued
    pending = None
    res_d = res_h.to(dev)
    rel_d = torch.empty(B, dtype=torch.float64, device=dev)
    act_prev, act_d = None, None
    # reliable updates: the relative residual at each right-hand side's last
    # recomputed residual, and those due for one (replaced at the next
    # iteration's recomputation point, so no extra host wait)
    rel_ref = rel.clone() if (refine_drop and rtol is not None) else None
    due = torch.zeros(B, dtype=torch.bool)
    pending_ref = None
    while bool(act.any()):
        idx = None if (bool(act.all()) or not COMPACT_ACTIVE) else \
            [i for i in range(B) if bool(act[i])]
        if act_prev is None or not torch.equal(act, act_prev):
            act_prev, act_d = act.clone(), act.to(dev)   # uploaded when it changes
        Ap = prob.A_active(p_op, idx)                         # :82
        # :84-95 in one call (alpha, x, r, z, dot(r, z), beta, p, res), with
        # sqrt(dot(r, r) / dot(b, b)) for the relative test
        _cg_step(prob, x, r, z, p, Ap, pinv_arr, p_op, res_d, act_d,
                 bb_d if rtol is not None else None, rel_d if rtol is not None else None, B)
        rel_cur = rel_d
        if pending_ref is not None:
            # the last read's reliable-update bookkeeping, done while the
            # device runs this iteration
            rel_p, tr_p, act_p = pending_ref
            pending_ref = None
            rel_ref = torch.where(torch.tensor(tr_p), rel_p, rel_ref)
            due = act_p & (rel_p < rel_ref * refine_drop) & (rel_p >= rtol)
            if bool(due.any()):
                # batched: with one right-hand side due, every active one at
                # least halfway there (in decades) is refined with it
                due = act_p & (rel_p < rel_ref * refine_drop ** 0.5) & (rel_p >= rtol)
        for i in (range(B) if idx is None else idx):
            true_r[i] = False
        k += act.long()
        it += 1
        ref = act.clone() if (recompute_every and it % recompute_every == 0) else None
        if rel_ref is not None and bool((due & act).any()):
            ref = (due & act) if ref is None else (ref | (due & act))
        if ref is not None:
            ridx = None if bool(ref.all()) else [i for i in range(B) if bool(ref[i])]
            ref_d = act_d if torch.equal(ref, act) else ref.to(dev)
            res_d = torch.where(ref_d, residual(x, ridx), res_d)
            rel_cur = rel_dev() if rtol is not None else None
        if pending is not None:
            record(*pending)
            pending = None
        upd = act
        if rtol is not None:
            res_h, rel = (yield from _host(torch.stack([res_d, rel_cur])))
            # before a right-hand side stops on the relative test, its
            # recursive residual is replaced by the recomputed b - A x (FP64
            # arithmetic), and it stops only if that one passes too.  The
            # recomputation is batched: a crossing right-hand side pauses
            # until half the batch waits or none is left iterating, and only
            # the waiting ones are recomputed (a paused right-hand side's
            # iterates do not change, so this only saves FP64 operator
            # applications).
            verify |= act & (rel < rtol)
            going = act & ~verify & (res_h >= tol) & (k < maxiter)
            if bool(verify.any()) and (not bool(going.any())
                                       or int(verify.sum()) >= max(1, B // 2)):
                vidx = [i for i in range(B) if bool(verify[i])]
                res_d = torch.where(verify.to(dev), residual(x, vidx), res_d)
                res_h, rel = (yield from _host(torch.stack([res_d, rel_dev()])))
                upd = act | verify
                verify[:] = False
            if rel_ref is not None:
                pending_ref = (rel, list(true_r), act)
        else:
            res_h = yield from _host(res_d)
        res_curr = torch.where(upd, res_h, res_curr)
        pending = (act, res_curr, rel)
        act = active_mask()
    if pending is not None:
        record(*pending)

    info.iterations = [int(v) for v in k]
    conv = res_curr < tol
    if rtol is not None:
        conv = conv | (rel < rtol)
    info.converged = [bool(v) for v in conv]
    if return_info:
        # the true residual of every right-hand side whose r is not already
        # a recomputed b - A x at its final x
        stale = [i for i in range(B) if not true_r[i]]
        if stale:
            residual(x, None if len(stale) == B else stale)
        rel_final = yield from rel_now()
        info.final_true_rel_residual = [float(v) for v in rel_final]
    out = unstack_matrices(x0, x.to(X0.dtype if X0.is_complex() else torch.complex128), lead)
    return (out, info) if return_info else out


def conjugate_gradient(x0, data, phi, mask, beam, noise_covariance, field_covariance, b_vector,
                       preconditioner, maxiter: int = 500, tol: float = 1e-1, *,
                       n: int = DEFAULT_NSTEPS, rtol: Optional[float] = None,
                       recompute_every: Optional[int] = None, return_info: bool = False,
                       refine_drop="auto"):
    """PCG of wiener_filter.py:51-105 (see conjugate_gradient_process)."""
    return run_process(conjugate_gradient_process(
        x0, data, phi, mask, beam, noise_covariance, field_covariance, b_vector, preconditioner,
        maxiter, tol, n=n, rtol=rtol, recompute_every=recompute_every, return_info=return_info,
        refine_drop=refine_drop))


def wiener_filter(field, phi, data, field_covariance, noise_covariance, mask, beam, *,
                  n: int = DEFAULT_NSTEPS, maxiter: int = 500, tol: float = 1e-1,
                  rtol: Optional[float] = None, recompute_every: Optional[int] = None,
                  return_info: bool = False, refine_drop="auto"):
    """wiener_filter.py:10-26: b = -gradf(0, d), P = Cf^-1 + B M Cn^-1 M B,
    PCG from `field`."""
    return run_process(wiener_filter_process(
        field, phi, data, field_covariance, noise_covariance, mask, beam, n=n, maxiter=maxiter,
        tol=tol, rtol=rtol, recompute_every=recompute_every, return_info=return_info,
        refine_drop=refine_drop))


def wiener_filter_process(field, phi, data, field_covariance, noise_covariance, mask, beam, *,
                          n: int = DEFAULT_NSTEPS, maxiter: int = 500, tol: float = 1e-1,
                          rtol: Optional[float] = None, recompute_every: Optional[int] = None,
                          return_info: bool = False, refine_drop="auto"):
    """wiener_filter as a process (see conjugate_gradient_process)."""
    b = get_b_vector(phi, data, field_covariance, noise_covariance, mask, beam, n=n, like=field)
    P = hessian_logpdf_preconditioner(field_covariance, beam, mask, noise_covariance)
    return (yield from conjugate_gradient_process(
        field, data, phi, mask, beam, noise_covariance, field_covariance, b, P, maxiter, tol, n=n,
        rtol=rtol, recompute_every=recompute_every, return_info=return_info,
        refine_drop=refine_drop))
