"""Iterative solvers, operators, preconditioners and field vectors (device
resident, through the C ABI).

Mirrors components/science/source/solver/sci_iterative_solver_mod.f90: every
solver type takes the reference constructor's arguments (lin_op, prec, r_tol,
a_tol, max_iter, monitor_convergence, fail_on_non_converged, and gcrk /
rho_relax / lmin, lmax where the type has them) and ``apply(x, b)`` solves
lin_op x = b in place:

* ``ConjugateGradient`` (:465-592), ``BiCGStab`` (:645-779), ``GMRES``
  (:830-1003), ``FGMRES`` (:1055-1237), ``GCR`` (:1288-1408),
  ``PreconditionOnly`` (:1664-1694), ``Jacobi`` (:1745-1815) and
  ``Chebyshev`` (:1868-1940);
* operators: an ``api.LMAOperator`` (mass_matrix_operator_type),
  ``BlockOperator`` / ``DenseOperator`` (dense_operator_type,
  sci_dense_operator_alg_mod.x90:147-182) and ``CallbackOperator`` (any
  abstract_linear_operator_type the caller writes);
* preconditioners: ``NullPreconditioner``, ``DiagonalPreconditioner``,
  ``ColumnwisePreconditioner`` (columnwise_op_appinv) and
  ``CallbackPreconditioner``;
* ``FieldVector``: field_vector_type / r_solver_field_vector_type with n
  fields (infrastructure/source/field/vector_mod.F90:17-190).

CG and GMRES on a plain LMA operator with the null or diagonal
preconditioner take the fused single-field path (lfr_cg, lfr_gmres); every
other combination runs lfr_iterative_solve.  ``RefinedCG`` wraps
lfr_refined_cg: FP64 residuals around FP32 (r_solver) CG corrections.
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


LFR_MAX_FIELDS = 8
(SOLVER_CG, SOLVER_BICGSTAB, SOLVER_GMRES, SOLVER_FGMRES, SOLVER_GCR, SOLVER_PRECONDITION_ONLY,
 SOLVER_JACOBI, SOLVER_CHEBYSHEV) = range(8)


class VectorLayout(C.Structure):
    _fields_ = [("prec", C.c_int), ("nfield", C.c_int), ("undf", C.c_int64 * LFR_MAX_FIELDS)]


class BlockEntry(C.Structure):
    _fields_ = [("row", C.c_int), ("col", C.c_int), ("matrix_d", _P), ("to_space", _P),
                ("from_space", _P), ("plan", _P), ("transpose", C.c_int)]


class IterativeSolverParams(C.Structure):
    _fields_ = [("method", C.c_int), ("r_tol", C.c_double), ("a_tol", C.c_double),
                ("max_iter", C.c_int), ("monitor_convergence", C.c_int),
                ("fail_on_non_converged", C.c_int), ("gcrk", C.c_int), ("rho_relax", C.c_double),
                ("lmin", C.c_double), ("lmax", C.c_double)]


APPLY_FN = C.CFUNCTYPE(C.c_int, _P, _P, _P, _P)


def _sig_generic():
    L = lib()
    LP = C.POINTER(VectorLayout)
    L.lfr_vector_size.argtypes = [LP]
    L.lfr_vector_size.restype = C.c_int
    L.lfr_vector_length.argtypes = [LP]
    L.lfr_vector_length.restype = C.c_int64
    L.lfr_vector_dot.argtypes = [_P, LP, _P, _P, C.POINTER(C.c_double)]
    L.lfr_vector_norm.argtypes = [_P, LP, _P, C.POINTER(C.c_double)]
    L.lfr_vector_field_norm.argtypes = [_P, LP, _P, C.c_int, C.POINTER(C.c_double)]
    L.lfr_linop_lma.argtypes = [_P, C.POINTER(LmaOp), C.POINTER(_P)]
    L.lfr_linop_block.argtypes = [_P, LP, C.c_int, C.POINTER(BlockEntry), C.POINTER(_P)]
    L.lfr_linop_callback.argtypes = [_P, LP, APPLY_FN, _P, C.POINTER(_P)]
    L.lfr_precon_null.argtypes = [_P, LP, C.POINTER(_P)]
    L.lfr_precon_diagonal.argtypes = [_P, LP, _P, C.POINTER(_P)]
    L.lfr_precon_columnwise.argtypes = [_P, LP, _P, C.c_int, _P, _P, C.c_int, C.POINTER(_P)]
    L.lfr_precon_callback.argtypes = [_P, LP, APPLY_FN, _P, C.POINTER(_P)]
    L.lfr_iterative_solve.argtypes = [_P, C.POINTER(IterativeSolverParams), _P, _P, _P, _P,
                                      C.POINTER(SolverStats)]
    for n in ("lfr_vector_dot", "lfr_vector_norm", "lfr_vector_field_norm", "lfr_linop_lma",
              "lfr_linop_block", "lfr_linop_callback", "lfr_precon_null", "lfr_precon_diagonal",
              "lfr_precon_columnwise", "lfr_precon_callback", "lfr_iterative_solve"):
        getattr(L, n).restype = C.c_int
    return L


def _layout(dtype, undf) -> VectorLayout:
    undf = [int(u) for u in undf]
    if not 1 <= len(undf) <= LFR_MAX_FIELDS:
        raise LfrError(1, f"a field vector holds 1..{LFR_MAX_FIELDS} fields, got {len(undf)}")
    lay = VectorLayout()
    lay.prec = 4 if dtype == torch.float32 else 8
    lay.nfield = len(undf)
    for i, u in enumerate(undf):
        lay.undf[i] = u
    return lay


class FieldVector:
    """field_vector_type (float64) / r_solver_field_vector_type (float32) of
    n fields, stored end to end in one device tensor ``data``
    (sci_field_vector_mod.x90:20-56).  ``field(i)`` is a view of field i
    (0-based); the elementwise methods act on the whole tensor, the
    reductions add per-field inner products in field order."""

    def __init__(self, ctx: "api.Context", undf, dtype=torch.float64, data: torch.Tensor | None = None):
        self.ctx = ctx
        self.undf = [int(u) for u in undf]
        n = sum(self.undf)
        if data is None:
            data = torch.zeros(n, dtype=dtype, device=ctx.device)
        if data.dim() != 1 or data.numel() != n or not data.is_contiguous() or data.device != ctx.device:
            raise LfrError(1, f"field vector data must be a contiguous 1-D tensor of {n} values on {ctx.device}")
        if data.dtype not in (torch.float32, torch.float64):
            raise LfrError(1, "field vectors are float32 or float64")
        self.data = data
        self.layout = _layout(data.dtype, self.undf)

    @classmethod
    def of(cls, ctx, fields):
        """A new vector holding copies of the given 1-D tensors, in order."""
        fields = list(fields)
        v = cls(ctx, [f.numel() for f in fields], fields[0].dtype)
        for i, f in enumerate(fields):
            v.field(i).copy_(f)
        return v

    def field(self, i: int) -> torch.Tensor:
        o = sum(self.undf[:i])
        return self.data[o:o + self.undf[i]]

    def vector_size(self) -> int:
        return _sig_generic().lfr_vector_size(C.byref(self.layout))

    def duplicate(self) -> "FieldVector":
        return FieldVector(self.ctx, self.undf, self.data.dtype)

    def _same(self, x: "FieldVector"):
        if not isinstance(x, FieldVector) or x.undf != self.undf or x.data.dtype != self.data.dtype:
            raise LfrError(1, "field vectors differ in fields or kind")

    def dot(self, x: "FieldVector") -> float:
        self._same(x)
        out = C.c_double()
        self.ctx.sync_stream()
        check(_sig_generic().lfr_vector_dot(self.ctx.h, C.byref(self.layout), _P(self.data.data_ptr()),
                                            _P(x.data.data_ptr()), C.byref(out)), self.ctx.h)
        return out.value

    def norm(self) -> float:
        out = C.c_double()
        self.ctx.sync_stream()
        check(_sig_generic().lfr_vector_norm(self.ctx.h, C.byref(self.layout), _P(self.data.data_ptr()),
                                             C.byref(out)), self.ctx.h)
        return out.value

    def field_norm(self, n: int) -> float:
        """Norm of field n (1-based, as the reference)."""
        out = C.c_double()
        self.ctx.sync_stream()
        check(_sig_generic().lfr_vector_field_norm(self.ctx.h, C.byref(self.layout),
                                                   _P(self.data.data_ptr()), int(n), C.byref(out)),
              self.ctx.h)
        return out.value

    def set_scalar(self, c: float):
        api.setval(self.ctx, self.data, c)

    def copy(self, source: "FieldVector"):
        self._same(source)
        api.copy(self.ctx, self.data, source.data)

    def axpy(self, alpha: float, x: "FieldVector"):
        """self = self + alpha x"""
        self._same(x)
        api.axpy(self.ctx, alpha, x.data, self.data)

    def aypx(self, alpha: float, x: "FieldVector"):
        """self = alpha self + x"""
        self._same(x)
        api.aypx(self.ctx, alpha, x.data, self.data)

    def axpby(self, a: float, x: "FieldVector", b: float, y: "FieldVector"):
        """self = a x + b y"""
        self._same(x)
        self._same(y)
        api.axpby(self.ctx, a, x.data, b, y.data, self.data)

    def scale(self, alpha: float):
        api.scale(self.ctx, alpha, self.data)


def _as_vector(ctx, t, layout_undf, dtype, name):
    if isinstance(t, FieldVector):
        if t.undf != list(layout_undf) or t.data.dtype != dtype:
            raise LfrError(1, f"{name}: field vector does not match the operator's fields")
        return t.data
    if not isinstance(t, torch.Tensor) or t.dim() != 1 or t.numel() != sum(layout_undf) \
            or t.dtype != dtype or not t.is_contiguous() or t.device != ctx.device:
        raise LfrError(1, f"{name} must be a FieldVector or a contiguous {dtype} tensor of "
                          f"{sum(layout_undf)} values on {ctx.device}")
    return t


class _DeviceArray:
    """A raw device buffer seen through __cuda_array_interface__ (callbacks)."""

    def __init__(self, ptr, n, dtype):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f4" if dtype == torch.float32 else "<f8",
                                         "data": (int(ptr), False), "version": 2, "strides": None}


def _wrap_callback(fn, ctx, undf, dtype):
    n = sum(undf)

    def cb(user, x, y, stream):
        try:
            xt = torch.as_tensor(_DeviceArray(x, n, dtype), device=ctx.device)
            yt = torch.as_tensor(_DeviceArray(y, n, dtype), device=ctx.device)
            fn(xt, yt)
            return 0
        except Exception as exc:   # reported as a nonzero status
            _wrap_callback.last_error = exc
            return 1
    return APPLY_FN(cb)


class LinearOperator:
    """abstract_linear_operator_type (sci_linear_operator_mod.f90:18-35)."""

    ctx = None
    undf: list = []
    dtype = torch.float64
    h = None

    def apply(self, x, y):
        """y = A x"""
        xd = _as_vector(self.ctx, x, self.undf, self.dtype, "x")
        yd = _as_vector(self.ctx, y, self.undf, self.dtype, "y")
        self.ctx.sync_stream()
        check(lib().lfr_linop_apply(self.ctx.h, self.h, _P(xd.data_ptr()), _P(yd.data_ptr())), self.ctx.h)
        return y

    def close(self):
        if getattr(self, "h", None):
            lib().lfr_linop_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MassMatrixOperator(LinearOperator):
    """mass_matrix_operator_type (sci_mass_matrix_operator_alg_mod.x90:72-119) over an
    api.LMAOperator (apply_bc and packed copies honoured)."""

    def __init__(self, op: "api.LMAOperator"):
        self.op, self.ctx = op, op.ctx
        self.undf, self.dtype = [op.to_space.undf], op.matrix.dtype
        self._a = lma_op(op)
        h = _P()
        check(_sig_generic().lfr_linop_lma(self.ctx.h, C.byref(self._a), C.byref(h)), self.ctx.h)
        self.h = h


class BlockOperator(LinearOperator):
    """y_row += A x_col for each entry (row, col, LMAOperator[, transpose]), y = 0 first.

    A continuous output uses matrix_vector (the operator's colouring), a
    discontinuous one (op.colouring None) dg_inc_matrix_vector, transpose=True
    transpose_matrix_vector (output on op.from_space)."""

    def __init__(self, fields, entries):
        self.spaces = list(fields)          # api.Space of each field
        ops = [e[2] for e in entries]
        self.ctx = ops[0].ctx
        self.dtype = ops[0].matrix.dtype
        self.undf = [sp.undf for sp in self.spaces]
        self.layout = _layout(self.dtype, self.undf)
        arr = (BlockEntry * len(entries))()
        self._keep = []
        for q, e in enumerate(entries):
            row, col, op = e[0], e[1], e[2]
            tr = bool(e[3]) if len(e) > 3 else False
            if op.matrix.dtype != self.dtype:
                raise LfrError(1, "block operator entries must share one kind")
            out_sp, in_sp = (op.from_space, op.to_space) if tr else (op.to_space, op.from_space)
            if out_sp is not self.spaces[row] or in_sp is not self.spaces[col]:
                raise LfrError(1, f"entry {q}: spaces do not match fields ({row}, {col})")
            plan = None
            if tr:
                if op.colouring is None:
                    raise LfrError(1, f"entry {q}: transpose needs the colouring of its output space")
                plan = out_sp.plan(op.colouring, 1)
            elif op.colouring is not None:
                plan = out_sp.plan(op.colouring, 0)
            self._keep.append(plan)
            arr[q] = BlockEntry(row, col, _P(op.matrix.data_ptr()), out_sp.h, in_sp.h,
                                plan.h if plan is not None else None, int(tr))
        self._ops = ops
        h = _P()
        check(_sig_generic().lfr_linop_block(self.ctx.h, C.byref(self.layout), len(entries), arr, C.byref(h)),
              self.ctx.h)
        self.h = h


class DenseOperator(BlockOperator):
    """dense_operator_type (sci_dense_operator_alg_mod.x90:147-182): y = 0;
    y1 += M_W0 x1 (matrix_vector); y2 += M_W3 x2 (dg_inc_matrix_vector)."""

    def __init__(self, mass_w0: "api.LMAOperator", mass_w3: "api.LMAOperator"):
        if mass_w3.colouring is not None:
            mass_w3 = api.LMAOperator(mass_w3.matrix, mass_w3.to_space, mass_w3.from_space, None)
        self.mass_w0, self.mass_w3 = mass_w0, mass_w3
        super().__init__([mass_w0.to_space, mass_w3.to_space], [(0, 0, mass_w0), (1, 1, mass_w3)])

    def extract_diagonal(self, diag: FieldVector) -> FieldVector:
        """diag = 0, then mm_diagonal of each mass matrix into its field (:190-202)."""
        diag.set_scalar(0.0)
        for i, op in enumerate((self.mass_w0, self.mass_w3)):
            plan = op.to_space.plan(op.colouring, 0) if op.colouring is not None else None
            check(lib().lfr_mm_diagonal(self.ctx.h, api._prec(op.matrix), _P(diag.field(i).data_ptr()),
                                        _P(op.matrix.data_ptr()), op.to_space.h,
                                        plan.h if plan is not None else None), self.ctx.h)
        return diag


class CallbackOperator(LinearOperator):
    """A caller-written operator: fn(x, y) sets y = A x on device tensors (queued on
    PyTorch's current stream)."""

    def __init__(self, ctx, undf, dtype, fn):
        self.ctx, self.undf, self.dtype = ctx, [int(u) for u in undf], dtype
        self.layout = _layout(dtype, self.undf)
        self._cb = _wrap_callback(fn, ctx, self.undf, dtype)
        h = _P()
        check(_sig_generic().lfr_linop_callback(ctx.h, C.byref(self.layout), self._cb, None, C.byref(h)),
              ctx.h)
        self.h = h


class _PrecHandle:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().lfr_precon_destroy(self.h)
        except Exception:
            pass


@dataclass
class ColumnwisePreconditioner:
    """y = columnwise_op_appinv(x): one Thomas solve per column (one-field vectors)."""
    cma: "api.CMAOperator"
    fp64_pivots: bool = False


class CallbackPreconditioner:
    """A caller-written preconditioner: fn(x, y) sets y = P^{-1} x."""

    def __init__(self, fn):
        self.fn = fn


def _precon_handle(prec, ctx, undf, dtype):
    L = _sig_generic()
    lay = _layout(dtype, undf)
    h = _P()
    keep = None
    if prec is None or isinstance(prec, NullPreconditioner):
        check(L.lfr_precon_null(ctx.h, C.byref(lay), C.byref(h)), ctx.h)
    elif isinstance(prec, DiagonalPreconditioner):
        d = prec.diag.data if isinstance(prec.diag, FieldVector) else prec.diag
        if d.dtype != dtype or d.numel() != sum(undf) or not d.is_contiguous():
            raise LfrError(1, "diagonal preconditioner: wrong kind or length")
        check(L.lfr_precon_diagonal(ctx.h, C.byref(lay), _P(d.data_ptr()), C.byref(h)), ctx.h)
        keep = d
    elif isinstance(prec, ColumnwisePreconditioner):
        cma = prec.cma
        if cma.matrix.dtype != dtype:
            raise LfrError(1, "column-solve preconditioner: wrong kind")
        check(L.lfr_precon_columnwise(ctx.h, C.byref(lay), _P(cma.matrix.data_ptr()), cma.nrow, cma.space.h,
                                      _P(cma.indirection.data_ptr()), int(prec.fp64_pivots), C.byref(h)), ctx.h)
        keep = cma
    elif isinstance(prec, CallbackPreconditioner):
        keep = _wrap_callback(prec.fn, ctx, undf, dtype)
        check(L.lfr_precon_callback(ctx.h, C.byref(lay), keep, None, C.byref(h)), ctx.h)
    else:
        raise LfrError(1, f"unsupported preconditioner {type(prec).__name__}")
    ph = _PrecHandle(h)
    ph.keep = keep
    return ph


def _linop(op):
    """A LinearOperator for the generic solvers (LMA operators are wrapped)."""
    if isinstance(op, LinearOperator):
        return op
    if isinstance(op, api.LMAOperator):
        return MassMatrixOperator(op)
    raise LfrError(1, f"unsupported operator {type(op).__name__}")


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


class _IterativeSolver:
    """abstract_iterative_solver_type (sci_iterative_solver_mod.f90:38-57): the
    generic path (lfr_iterative_solve) over any operator and preconditioner."""

    method = SOLVER_CG

    def __init__(self, lin_op, prec=None, r_tol: float = 1e-8, a_tol: float = 0.0, max_iter: int = 100,
                 monitor_convergence: bool = True, fail_on_non_converged: bool = False):
        self.lin_op = lin_op
        self.prec = prec or NullPreconditioner()
        self.r_tol, self.a_tol, self.max_iter = r_tol, a_tol, max_iter
        self.monitor_convergence = monitor_convergence
        self.fail_on_non_converged = fail_on_non_converged
        self.gcrk, self.rho_relax, self.lmin, self.lmax = 0, 0.0, 0.0, 0.0
        self.stats = None
        self.bytes = None

    def _generic(self, x, b) -> SolverStats:
        op = _linop(self.lin_op)
        ctx = op.ctx
        xd = _as_vector(ctx, x, op.undf, op.dtype, "x")
        bd = _as_vector(ctx, b, op.undf, op.dtype, "b")
        ph = _precon_handle(self.prec, ctx, op.undf, op.dtype)
        prm = IterativeSolverParams(self.method, self.r_tol, self.a_tol, self.max_iter,
                                    int(self.monitor_convergence), int(self.fail_on_non_converged),
                                    int(self.gcrk), float(self.rho_relax), float(self.lmin), float(self.lmax))
        st = SolverStats()
        ctx.sync_stream()
        rc = _sig_generic().lfr_iterative_solve(ctx.h, C.byref(prm), op.h, ph.h, _P(xd.data_ptr()),
                                                _P(bd.data_ptr()), C.byref(st))
        self.stats = st
        check(rc, ctx.h)
        nb = C.c_double(-1.0)
        check(lib().lfr_iterative_solve_bytes(ctx.h, C.byref(nb)), ctx.h)
        # algorithmic HBM bytes of this solve (None: a caller-written operator
        # or preconditioner took part)
        self.bytes = nb.value if nb.value >= 0 else None
        return st

    def apply(self, x, b) -> SolverStats:
        return self._generic(x, b)


def _fast_path(lin_op, prec) -> bool:
    return isinstance(lin_op, api.LMAOperator) and isinstance(prec, (NullPreconditioner, DiagonalPreconditioner)) \
        and not (isinstance(prec, DiagonalPreconditioner) and isinstance(prec.diag, FieldVector))


class ConjugateGradient(_IterativeSolver):
    """conjugate_gradient_type (:465-592)."""

    method = SOLVER_CG

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

    def apply(self, x, b, x_is_zero: bool = False) -> SolverStats:
        if not _fast_path(self.lin_op, self.prec):
            if x_is_zero:
                (x.data if isinstance(x, FieldVector) else x).zero_()
            return self._generic(x, b)
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

    method = SOLVER_GMRES

    def __init__(self, lin_op, prec=None, gcrk: int = 10, r_tol: float = 1e-8, a_tol: float = 0.0,
                 max_iter: int = 10, monitor_convergence: bool = True,
                 fail_on_non_converged: bool = False):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence,
                         fail_on_non_converged)
        self.gcrk = gcrk

    def apply(self, x, b, x_is_zero: bool = False):
        if not _fast_path(self.lin_op, self.prec):
            if x_is_zero:
                (x.data if isinstance(x, FieldVector) else x).zero_()
            return self._generic(x, b)
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


class BiCGStab(_IterativeSolver):
    """BiCGstab_type (sci_iterative_solver_mod.f90:645-779)."""

    method = SOLVER_BICGSTAB


class FGMRES(_IterativeSolver):
    """fgmres_type (:1055-1237): flexible GMRES(gcrk), right preconditioned; the
    correction starts from zero and is added to x at the end."""

    method = SOLVER_FGMRES

    def __init__(self, lin_op, prec=None, gcrk: int = 10, r_tol: float = 1e-8, a_tol: float = 0.0,
                 max_iter: int = 10, monitor_convergence: bool = True, fail_on_non_converged: bool = False):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, fail_on_non_converged)
        self.gcrk = gcrk


class GCR(FGMRES):
    """gcr_type (:1288-1408): restarted GCR(gcrk); the correction starts from
    P^{-1} b and is added to x at the end."""

    method = SOLVER_GCR


class PreconditionOnly(_IterativeSolver):
    """precondition_only_type (:1664-1694): x = P^{-1} b."""

    method = SOLVER_PRECONDITION_ONLY

    def __init__(self, lin_op, prec=None, monitor_convergence: bool = False):
        super().__init__(lin_op, prec, 0.0, 0.0, 1, monitor_convergence, False)


class Jacobi(_IterativeSolver):
    """jacobi_type (:1745-1815): x -= rho_relax P^{-1} (A x - b)."""

    method = SOLVER_JACOBI

    def __init__(self, lin_op, prec=None, r_tol: float = 1e-8, a_tol: float = 0.0, max_iter: int = 100,
                 monitor_convergence: bool = True, fail_on_non_converged: bool = False, rho_relax: float = 1.0):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, fail_on_non_converged)
        self.rho_relax = rho_relax


class Chebyshev(_IterativeSolver):
    """chebyshev_type (:1868-1940): max_iter Chebyshev steps from x = 0 with the
    eigenvalue bounds lmin, lmax of P^{-1} A."""

    method = SOLVER_CHEBYSHEV

    def __init__(self, lin_op, prec=None, r_tol: float = 1e-8, a_tol: float = 0.0, max_iter: int = 10,
                 monitor_convergence: bool = True, fail_on_non_converged: bool = False, lmin: float = 0.0,
                 lmax: float = 1.0):
        super().__init__(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, fail_on_non_converged)
        self.lmin, self.lmax = lmin, lmax
