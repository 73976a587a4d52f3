"""Differentiable operator applies and CG solves (torch.autograd front end).

All arithmetic runs in liblfr.so through the C ABI; PyTorch provides the
tensors and the autograd graph.

* ``lma_apply(M, x, op)``        y = M x  (matrix_vector overwrite / dg_matrix_vector)
* ``lma_apply_transpose(M, p, op)``  b = M^T p  (transpose_matrix_vector)
* ``cg_solve(M, b, op, iters)``   u = x_K of K unmonitored CG iterations from 0
                                  (conjugate_gradient_type%cg_solve,
                                  components/science/source/solver/sci_iterative_solver_mod.f90:495-592)

Backward passes:

* apply: xbar = M^T ybar (the transpose kernel is the adjoint of the apply);
  Mbar(ik,i,j) = ybar(map_to(i)+k) * x(map_from(j)+k) (``lfr_lma_outer``).
* CG: exact reverse mode of the K-iteration algorithm (alpha and beta depend
  on b, so the map b -> x_K is nonlinear and its derivative is not the
  linear-solve adjoint unless CG has converged).  Iterates are not stored:
  the forward pass keeps checkpoints (r, p, rz) every ``C = ceil(sqrt(K))``
  iterations; the reverse sweep recomputes one segment at a time from its
  checkpoint, then walks it backwards.  Memory is O(sqrt(K)) vectors; cost is
  one extra forward pass.
"""

from __future__ import annotations

import math

import torch

from . import api


def _zeros(space: "api.Space", dtype):
    return torch.zeros(space.undf, dtype=dtype, device=space.ctx.device)


def _apply(M: torch.Tensor, x: torch.Tensor, op: "api.LMAOperator") -> torch.Tensor:
    opm = op.with_matrix(M)
    y = torch.empty(op.to_space.undf, dtype=M.dtype, device=M.device)
    if op.to_space.fs.is_discontinuous:
        api.dg_matrix_vector(y, x, opm)
    else:
        api.matrix_vector(y, x, opm, overwrite=True)
    return y


def _apply_t(M: torch.Tensor, p: torch.Tensor, op: "api.LMAOperator") -> torch.Tensor:
    opm = op.with_matrix(M)
    b = torch.empty(op.from_space.undf, dtype=M.dtype, device=M.device)
    api.transpose_matrix_vector(b, p, opm, overwrite=True)
    return b


def _outer(ybar: torch.Tensor, x: torch.Tensor, op: "api.LMAOperator", out=None, alpha=1.0):
    """Mbar (+)= alpha * ybar (x) x, per cell-layer (lfr_lma_outer)."""
    from ._ffi import check, lib
    import ctypes as C
    acc = out is not None
    if out is None:
        out = torch.empty_like(op.matrix)
    ctx = op.ctx
    ctx.sync_stream()
    check(lib().lfr_lma_outer(ctx.h, api._prec(out), C.c_void_p(out.data_ptr()),
                              C.c_void_p(ybar.data_ptr()), C.c_void_p(x.data_ptr()),
                              op.to_space.h, op.from_space.h, float(alpha), int(acc)), ctx.h)
    return out


class _LMAApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, M, x, op):
        ctx.op = op
        ctx.save_for_backward(M, x)
        return _apply(M, x, op)

    @staticmethod
    def backward(ctx, ybar):
        M, x = ctx.saved_tensors
        op = ctx.op
        ybar = ybar.contiguous()
        xbar = _apply_t(M, ybar, op) if ctx.needs_input_grad[1] else None
        Mbar = _outer(ybar, x, op) if ctx.needs_input_grad[0] else None
        return Mbar, xbar, None


class _LMAApplyT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, M, p, op):
        ctx.op = op
        ctx.save_for_backward(M, p)
        return _apply_t(M, p, op)

    @staticmethod
    def backward(ctx, bbar):
        M, p = ctx.saved_tensors
        op = ctx.op
        bbar = bbar.contiguous()
        pbar = _apply(M, bbar, op) if ctx.needs_input_grad[1] else None
        Mbar = _outer(p, bbar, op) if ctx.needs_input_grad[0] else None
        return Mbar, pbar, None


def lma_apply(M: torch.Tensor, x: torch.Tensor, op: "api.LMAOperator") -> torch.Tensor:
    """y = M x, differentiable in M and x."""
    return _LMAApply.apply(M, x, op)


def lma_apply_transpose(M: torch.Tensor, p: torch.Tensor, op: "api.LMAOperator") -> torch.Tensor:
    """b = M^T p (op maps from_space -> to_space; b lives on from_space)."""
    return _LMAApplyT.apply(M, p, op)


# ---------------------------------------------------------------------------
# K-iteration CG with checkpointed reverse mode
# ---------------------------------------------------------------------------

class _CGRecorder:
    """Forward CG from x0 = 0 in the reference's statement order, using the
    library's apply, vector and FP64 inner-product entry points."""

    def __init__(self, M, op):
        self.M, self.op, self.ctx = M, op, op.ctx

    def start(self, b):
        ctx = self.ctx
        r = b.clone()                       # r = b - A*0 = b
        p = r.clone()                       # null preconditioner: z = r, p = z
        rz = api.dot(ctx, r, r)
        return r, p, rz

    def step(self, x, r, p, rz):
        """One iteration; returns (q, pq, alpha, rz_new, beta) and updates x, r, p in place."""
        ctx = self.ctx
        q = _apply(self.M, p, self.op)
        pq = api.dot(ctx, p, q)
        alpha = rz / pq
        if x is not None:
            api.axpy(ctx, alpha, p, x)
        api.axpy(ctx, -alpha, q, r)
        rz_new = api.dot(ctx, r, r)
        beta = rz_new / rz
        api.aypx(ctx, beta, r, p)           # p = beta p + r
        return q, pq, alpha, rz_new, beta


class _CGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, M, b, op, iters, seg):
        rec = _CGRecorder(M, op)
        x = torch.zeros_like(b)
        r, p, rz = rec.start(b)
        ckpt = {0: (r.clone(), p.clone(), rz)}
        for i in range(iters):
            rz = rec.step(x, r, p, rz)[3]
            if (i + 1) % seg == 0 and i + 1 < iters:
                ckpt[i + 1] = (r.clone(), p.clone(), rz)
        ctx.op, ctx.iters, ctx.seg, ctx.ckpt = op, iters, seg, ckpt
        ctx.save_for_backward(M, b)
        return x

    @staticmethod
    def backward(ctx, xbar):
        M, b = ctx.saved_tensors
        op, K, seg, ckpt = ctx.op, ctx.iters, ctx.seg, ctx.ckpt
        actx = op.ctx
        rec = _CGRecorder(M, op)
        need_M = ctx.needs_input_grad[0]
        xbar = xbar.contiguous().clone()          # constant through the sweep
        rbar = torch.zeros_like(b)
        pbar = torch.zeros_like(b)
        rzbar = 0.0                               # adjoint of rz_{i+1} carried down
        Mbar = torch.zeros_like(M) if need_M else None
        starts = sorted(ckpt)
        for s0 in reversed(starts):
            s1 = min(s0 + seg, K)
            # recompute the segment from its checkpoint
            r, p, rz = (t.clone() if torch.is_tensor(t) else t for t in ckpt[s0])
            hist = []
            for i in range(s0, s1):
                p_i = p.clone()
                q, pq, alpha, rz_new, beta = rec.step(None, r, p, rz)
                hist.append((p_i, q, r.clone(), pq, alpha, rz, rz_new, beta))
                rz = rz_new
            for (p_i, q_i, r_next, pq, alpha, rz_i, rz_next, beta) in reversed(hist):
                # p_{i+1} = r_{i+1} + beta_i p_i
                api.axpy(actx, 1.0, pbar, rbar)
                betabar = api.dot(actx, pbar, p_i)
                api.scale(actx, beta, pbar)
                # beta_i = rz_{i+1} / rz_i
                rzbar_next = rzbar + betabar / rz_i
                rzbar_i = -betabar * rz_next / (rz_i * rz_i)
                # rz_{i+1} = <r_{i+1}, r_{i+1}>
                api.axpy(actx, 2.0 * rzbar_next, r_next, rbar)
                # r_{i+1} = r_i - alpha q_i ; x_{i+1} = x_i + alpha p_i
                alphabar = -api.dot(actx, rbar, q_i) + api.dot(actx, xbar, p_i)
                qbar = torch.zeros_like(b)
                api.axpy(actx, -alpha, rbar, qbar)
                api.axpy(actx, alpha, xbar, pbar)
                # alpha_i = rz_i / pq_i
                rzbar_i += alphabar / pq
                pqbar = -alphabar * rz_i / (pq * pq)
                # pq_i = <p_i, q_i>
                api.axpy(actx, pqbar, q_i, pbar)
                api.axpy(actx, pqbar, p_i, qbar)
                # q_i = A p_i
                api.axpy(actx, 1.0, _apply_t(M, qbar, op), pbar)
                if need_M:
                    _outer(qbar, p_i, op, out=Mbar)
                rzbar = rzbar_i
        # p_0 = r_0, rz_0 = <r_0, r_0>, r_0 = b
        api.axpy(actx, 1.0, pbar, rbar)
        api.axpy(actx, 2.0 * rzbar, b, rbar)
        return Mbar, rbar, None, None, None


def cg_solve(M: torch.Tensor, b: torch.Tensor, op: "api.LMAOperator", iters: int = 20,
             segment: int | None = None) -> torch.Tensor:
    """x_K of K CG iterations on M x = b from x = 0, differentiable in M and b."""
    if op.to_space is not op.from_space:
        raise ValueError("cg_solve needs a square operator")
    seg = segment or max(1, math.ceil(math.sqrt(iters)))
    return _CGSolve.apply(M, b, op, iters, seg)
