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
  linear-solve adjoint unless CG has converged), on the device
  (lfr_cg_tape_*).  The iterates r, p are not stored: the forward keeps them
  every ``C = ceil(sqrt(K))`` iterations plus the operator products A p_i
  and the scalars; the reverse sweep recomputes each segment's iterates from
  its checkpoint by vector updates alone and walks it backwards, one
  transpose apply per step, with no host synchronisation.
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
# K-iteration CG with checkpointed reverse mode (lfr_cg_tape_*)
# ---------------------------------------------------------------------------

class CGTape:
    """A device tape of K CG iterations on op (lfr_cg_tape_create): forward
    records the scalars and operator products, reverse recomputes the
    iterates segment by segment and returns the gradients."""

    def __init__(self, M: torch.Tensor, op: "api.LMAOperator", iters: int, segment: int = 0,
                 record: bool = True, packed: "api.PackedOperator | None" = None,
                 packed_t: "api.PackedOperator | None" = None):
        from ._ffi import check, lib
        from .solvers import LmaOp
        import ctypes as C
        L = lib()
        L.lfr_cg_tape_create.argtypes = [C.c_void_p, C.POINTER(LmaOp), C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                         C.c_int, C.POINTER(C.c_void_p)]
        L.lfr_cg_tape_create.restype = C.c_int
        sp = op.to_space
        self.ctx, self.op, self.M = op.ctx, op, M
        plan = sp.plan(op.colouring, 0)
        plan_t = sp.plan(op.colouring, 1)
        # optional packed copies of M (kind 0) and of its transpose (kind 1)
        self._a = LmaOp(api._prec(M), C.c_void_p(M.data_ptr()), sp.h, plan.h, 0,
                        packed.h if packed is not None else None)
        self._keep = (plan, plan_t, packed, packed_t)
        h = C.c_void_p()
        self.ctx.sync_stream()
        check(L.lfr_cg_tape_create(self.ctx.h, C.byref(self._a), plan_t.h,
                                   packed_t.h if packed_t is not None else None, int(iters), int(segment),
                                   int(record), C.byref(h)), self.ctx.h)
        self.h = h

    def forward(self, b: torch.Tensor) -> torch.Tensor:
        from ._ffi import check, lib
        import ctypes as C
        x = torch.empty_like(b)
        self.ctx.sync_stream()
        check(lib().lfr_cg_tape_forward(self.ctx.h, self.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr())),
              self.ctx.h)
        return x

    def reverse(self, b: torch.Tensor, xbar: torch.Tensor, want_M: bool):
        from ._ffi import check, lib
        import ctypes as C
        bbar = torch.empty_like(b)
        Mbar = torch.empty_like(self.M) if want_M else None
        self.ctx.sync_stream()
        check(lib().lfr_cg_tape_reverse(self.ctx.h, self.h, C.c_void_p(b.data_ptr()), C.c_void_p(xbar.data_ptr()),
                                        C.c_void_p(bbar.data_ptr()),
                                        C.c_void_p(Mbar.data_ptr()) if want_M else None, 0), self.ctx.h)
        return bbar, Mbar

    def close(self):
        from ._ffi import lib
        if getattr(self, "h", None):
            lib().lfr_cg_tape_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _CGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, M, b, op, iters, seg):
        record = bool(ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
        tape = CGTape(M, op, iters, seg, record)
        x = tape.forward(b.contiguous())
        if record:
            ctx.tape = tape
            ctx.save_for_backward(M, b)
        else:
            tape.close()
        return x

    @staticmethod
    def backward(ctx, xbar):
        M, b = ctx.saved_tensors
        bbar, Mbar = ctx.tape.reverse(b.contiguous(), xbar.contiguous(), bool(ctx.needs_input_grad[0]))
        ctx.tape.close()
        ctx.tape = None
        return Mbar, bbar, None, None, None


def cg_solve(M: torch.Tensor, b: torch.Tensor, op: "api.LMAOperator", iters: int = 20,
             segment: int | None = None) -> torch.Tensor:
    """x_K of K CG iterations on M x = b from x = 0, differentiable in M and b."""
    if op.to_space is not op.from_space:
        raise ValueError("cg_solve needs a square operator")
    seg = segment or max(1, math.ceil(math.sqrt(iters)))
    return _CGSolve.apply(M, b, op, iters, seg)
