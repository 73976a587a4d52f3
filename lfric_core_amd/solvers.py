"""Iterative solvers over LMA operators (device resident, through the C ABI).

``ConjugateGradient`` mirrors conjugate_gradient_type
(components/science/source/solver/sci_iterative_solver_mod.f90:465-592):
same constructor arguments (lin_op, prec, r_tol, a_tol, max_iter,
monitor_convergence, fail_on_non_converged) and ``apply(x, b)`` solving
lin_op x = b in place.  ``RefinedCG`` wraps lfr_refined_cg: FP64 residuals
around FP32 (r_solver) CG corrections.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

import torch

from . import api
from ._ffi import LfrError, check, lib

_P = C.c_void_p


class LmaOp(C.Structure):
    _fields_ = [("prec", C.c_int), ("matrix_d", _P), ("space", _P), ("plan", _P), ("apply_bc", C.c_int),
                ("packed", _P)]


class SolverParams(C.Structure):
    _fields_ = [("r_tol", C.c_double), ("a_tol", C.c_double), ("max_iter", C.c_int),
                ("monitor_convergence", C.c_int), ("fail_on_non_converged", C.c_int),
                ("preconditioner", C.c_int), ("diag_d", _P), ("gcrk", C.c_int),
                ("x_is_zero", C.c_int)]


class SolverStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("r_nrm0", C.c_double),
                ("r_nrm", C.c_double)]


class RefineStats(C.Structure):
    _fields_ = [("outer", C.c_int), ("inner_total", C.c_int), ("converged", C.c_int),
                ("b_nrm", C.c_double), ("r_nrm", C.c_double)]


def _sig():
    L = lib()
    L.lfr_cg.argtypes = [_P, C.POINTER(LmaOp), C.POINTER(SolverParams), _P, _P,
                         C.POINTER(SolverStats)]
    L.lfr_cg.restype = C.c_int
    L.lfr_refined_cg.argtypes = [_P, C.POINTER(LmaOp), C.POINTER(LmaOp), C.c_double, C.c_int,
                                 C.POINTER(SolverParams), _P, _P, C.POINTER(RefineStats)]
    L.lfr_refined_cg.restype = C.c_int
    L.lfr_gmres.argtypes = L.lfr_cg.argtypes
    L.lfr_gmres.restype = C.c_int
    L.lfr_refined_cg_multi.argtypes = [_P, _P, C.POINTER(LmaOp), C.c_double, C.c_int, C.c_double,
                                       C.c_int, _P, _P, C.c_int, C.POINTER(RefineStats), _P]
    L.lfr_refined_cg_multi.restype = C.c_int
    return L


def lma_op(op: "api.LMAOperator") -> LmaOp:
    if op.to_space is not op.from_space:
        raise LfrError(1, "solvers need a square operator (to_space is from_space)")
    plan = op.to_space.plan(op.colouring, 0)
    pk = op.packed
    if pk is not None and (pk.kind != 0 or pk.plan is not plan):
        raise LfrError(1, "op.packed must be a kind-0 pack of this operator with its colouring")
    return LmaOp(api._prec(op.matrix), _P(op.matrix.data_ptr()), op.to_space.h, plan.h,
                 int(bool(op.apply_bc)), pk.h if pk is not None else None)


@dataclass
class NullPreconditioner:
    """null_preconditioner_type: y = x (sci_null_preconditioner_alg_mod.x90:73-95)."""


@dataclass
class DiagonalPreconditioner:
    """diagonal_preconditioner_type: y = x / diag (sci_diagonal_preconditioner_alg_mod.x90:71-106)."""
    diag: torch.Tensor


class ConjugateGradient:
    def __init__(self, lin_op: "api.LMAOperator", prec=None, r_tol: float = 1e-8,
                 a_tol: float = 0.0, max_iter: int = 100, monitor_convergence: bool = True,
                 fail_on_non_converged: bool = False):
        self.lin_op = lin_op
        self.prec = prec or NullPreconditioner()
        self.r_tol, self.a_tol, self.max_iter = r_tol, a_tol, max_iter
        self.monitor_convergence = monitor_convergence
        self.fail_on_non_converged = fail_on_non_converged
        self.stats = None

    def params(self, x_is_zero=False) -> SolverParams:
        diag = None
        pk = 0
        if isinstance(self.prec, DiagonalPreconditioner):
            pk = 1
            d = self.prec.diag
            if d.dtype != self.lin_op.matrix.dtype or d.numel() != self.lin_op.to_space.undf:
                raise LfrError(1, "diagonal preconditioner: wrong dtype or length")
            diag = _P(d.data_ptr())
        return SolverParams(self.r_tol, self.a_tol, self.max_iter, int(self.monitor_convergence),
                            int(self.fail_on_non_converged), pk, diag, 0, int(x_is_zero))

    def apply(self, x: torch.Tensor, b: torch.Tensor, x_is_zero: bool = False) -> SolverStats:
        op = self.lin_op
        api._field(x, op.to_space, "x", op.matrix.dtype)
        api._field(b, op.to_space, "b", op.matrix.dtype)
        L = _sig()
        ctx = op.ctx
        ctx.sync_stream()
        A = lma_op(op)
        P = self.params(x_is_zero)
        st = SolverStats()
        check(L.lfr_cg(ctx.h, C.byref(A), C.byref(P), _P(x.data_ptr()), _P(b.data_ptr()),
                       C.byref(st)), ctx.h)
        self.stats = st
        return st


class RefinedCG:
    """FP32 CG corrections inside an FP64 residual loop, to ||b - A x|| <= r_tol ||b||."""

    def __init__(self, op32: "api.LMAOperator", op64: "api.LMAOperator", r_tol: float = 1e-10,
                 max_outer: int = 20, inner_r_tol: float = 1e-5, inner_max_iter: int = 200):
        self.op32, self.op64 = op32, op64
        self.r_tol, self.max_outer = r_tol, max_outer
        self.inner = SolverParams(inner_r_tol, 0.0, inner_max_iter, 1, 0, 0, None, 0, 1)
        self.stats = None

    def apply(self, x64: torch.Tensor, b64: torch.Tensor) -> RefineStats:
        api._field(x64, self.op64.to_space, "x", torch.float64)
        api._field(b64, self.op64.to_space, "b", torch.float64)
        L = _sig()
        ctx = self.op64.ctx
        ctx.sync_stream()
        A32, A64 = lma_op(self.op32), lma_op(self.op64)
        st = RefineStats()
        check(L.lfr_refined_cg(ctx.h, C.byref(A32), C.byref(A64), self.r_tol, self.max_outer,
                               C.byref(self.inner), _P(x64.data_ptr()), _P(b64.data_ptr()),
                               C.byref(st)), ctx.h)
        self.stats = st
        return st


class RefinedCGMulti:
    """Sixteen simultaneous refined solves: FP32 CG on the MFMA multi-field
    operator inside an FP64 residual loop (lfr_refined_cg_multi)."""

    def __init__(self, mop: "api.MultiOperator", op64: "api.LMAOperator", r_tol: float = 1e-10,
                 max_outer: int = 20, inner_r_tol: float = 1e-5, inner_max_iter: int = 500):
        self.mop, self.op64 = mop, op64
        self.r_tol, self.max_outer = r_tol, max_outer
        self.inner_r_tol, self.inner_max_iter = inner_r_tol, inner_max_iter
        self.stats = None
        self.rel_res = None

    def apply(self, x64: torch.Tensor, b64: torch.Tensor) -> RefineStats:
        shape = (self.op64.to_space.undf, 16)
        for t, nm in ((x64, "x"), (b64, "b")):
            if tuple(t.shape) != shape or t.dtype != torch.float64 or not t.is_contiguous() \
                    or t.device != self.op64.ctx.device:
                raise LfrError(1, f"{nm} must be a contiguous float64 tensor of shape {shape}")
        L = _sig()
        ctx = self.op64.ctx
        ctx.sync_stream()
        A64 = lma_op(self.op64)
        st = RefineStats()
        rel = np.zeros(16)
        check(L.lfr_refined_cg_multi(ctx.h, self.mop.h, C.byref(A64), self.r_tol, self.max_outer,
                                     self.inner_r_tol, self.inner_max_iter, _P(x64.data_ptr()),
                                     _P(b64.data_ptr()), 16, C.byref(st),
                                     rel.ctypes.data_as(_P)), ctx.h)
        self.stats, self.rel_res = st, rel
        return st


class GMRES(ConjugateGradient):
    """gmres_type (sci_iterative_solver_mod.f90:795-1003): restarted GMRES(gcrk);
    max_iter counts restarts."""

    def __init__(self, lin_op, prec=None, gcrk: int = 10, r_tol: float = 1e-8, a_tol: float = 0.0,
                 max_iter: int = 10, monitor_convergence: bool = True,
                 fail_on_non_converged: bool = False):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence,
                         fail_on_non_converged)
        self.gcrk = gcrk

    def apply(self, x, b, x_is_zero: bool = False):
        op = self.lin_op
        api._field(x, op.to_space, "x", op.matrix.dtype)
        api._field(b, op.to_space, "b", op.matrix.dtype)
        if x_is_zero:
            x.zero_()
        L = _sig()
        ctx = op.ctx
        ctx.sync_stream()
        A = lma_op(op)
        P = self.params(False)
        P.gcrk = self.gcrk
        st = SolverStats()
        check(L.lfr_gmres(ctx.h, C.byref(A), C.byref(P), _P(x.data_ptr()), _P(b.data_ptr()),
                          C.byref(st)), ctx.h)
        self.stats = st
        return st
