"""Python front end over the C ABI (include/lfr.h), on PyTorch device tensors.

Mirrors the reference's invoke-level vocabulary:

* ``matrix_vector(lhs, x, op)``            invoke(matrix_vector_kernel_type(lhs, x, op))
* ``dg_matrix_vector(lhs, x, op)``         invoke(dg_matrix_vector_kernel_type(lhs, x, op))
* ``dg_inc_matrix_vector(lhs, x, op)``     invoke(dg_inc_matrix_vector_kernel_type(lhs, x, op))
* ``transpose_matrix_vector(lhs, x, op)``  invoke(transpose_matrix_vector_kernel_type(lhs, x, op))

Fields are 1-D tensors of length undf (float32 = r_solver / r_single,
float64 = r_def / r_double) on the context's device; operators are
``LMAOperator`` objects holding local_stencil(ncell_3d, ndf_to, ndf_from) as a
tensor of shape (ndf_from, ndf_to, ncell_3d) (the same memory as the Fortran
array).  Every argument is checked on the host before a kernel is queued.
PyTorch supplies device memory and the stream; all arithmetic is in liblfr.so.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from ._ffi import LfrError, check, lib
from .function_space import Colouring, FunctionSpace

_P = C.c_void_p


def _prec(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return 4
    if t.dtype == torch.float64:
        return 8
    raise LfrError(1, f"unsupported dtype {t.dtype}; fields are float32 or float64")


class Context:
    """A device context (lfr_init); work is queued on PyTorch's current stream."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise LfrError(2, "no GPU: liblfr needs a HIP device")
        self.device = torch.device("cuda", device)
        h = _P()
        rc = lib().lfr_init(device, C.byref(h))
        if rc != 0:
            raise LfrError(rc, "lfr_init failed (see stderr)")
        self.h = h

    def sync_stream(self) -> None:
        """Queue subsequent library work on PyTorch's current stream."""
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().lfr_set_stream(self.h, _P(s)), self.h)

    def synchronize(self) -> None:
        check(lib().lfr_synchronize(self.h), self.h)

    # -- autotuner: launch-shape choices per (kernel family, shape) key ------
    def tune_entries(self) -> dict:
        """Every stored launch-shape choice, key -> variant index."""
        n = C.c_int()
        check(lib().lfr_tune_count(self.h, C.byref(n)), self.h)
        out = {}
        buf = C.create_string_buffer(256)
        choice = C.c_int()
        for i in range(n.value):
            check(lib().lfr_tune_entry(self.h, i, buf, 256, C.byref(choice)), self.h)
            out[buf.value.decode()] = choice.value
        return out

    def tune_set(self, key: str, choice: int) -> None:
        check(lib().lfr_tune_set(self.h, key.encode(), int(choice)), self.h)

    def tune_get(self, key: str) -> int:
        choice = C.c_int()
        check(lib().lfr_tune_get(self.h, key.encode(), C.byref(choice)), self.h)
        return choice.value

    def tune_clear(self) -> None:
        check(lib().lfr_tune_clear(self.h), self.h)

    def tune_save(self, path) -> None:
        check(lib().lfr_tune_save(self.h, str(path).encode()), self.h)

    def tune_load(self, path) -> None:
        check(lib().lfr_tune_load(self.h, str(path).encode()), self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().lfr_finalise(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Space:
    """A function space's dofmap on the device (lfr_space_create)."""

    def __init__(self, ctx: Context, fs: FunctionSpace):
        self.ctx, self.fs = ctx, fs
        dm = np.ascontiguousarray(fs.dofmap, dtype=np.int32)
        h = _P()
        check(lib().lfr_space_create(ctx.h, dm.ctypes.data_as(_P), fs.ndf, fs.ncells,
                                     fs.nlayers, fs.undf, C.byref(h)), ctx.h)
        self.h = h
        self._plans = {}

    @property
    def undf(self) -> int:
        return self.fs.undf

    @property
    def ndf(self) -> int:
        return self.fs.ndf

    def set_boundary(self, boundary=None) -> None:
        """Attach boundary_value(ndf, 2) (default: function_space.boundary_dofs)."""
        from .function_space import boundary_dofs
        bv = boundary_dofs(self.fs) if boundary is None else np.asarray(boundary, dtype=np.int32)
        if bv.shape != (self.ndf, 2):
            raise LfrError(1, f"boundary values must have shape {(self.ndf, 2)}")
        self._bv = np.asfortranarray(bv, dtype=np.int32)
        check(lib().lfr_space_set_boundary(self.h, self._bv.ctypes.data_as(_P)), self.ctx.h)

    def plan(self, colouring: Colouring, kind: int) -> "Plan":
        key = (id(colouring), kind)
        if key not in self._plans:
            self._plans[key] = Plan(self, colouring, kind)
        return self._plans[key]

    def __del__(self):
        try:
            for p in self._plans.values():
                p.close()
            if self.h:
                lib().lfr_space_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Plan:
    """Colour-ordered gather plan of a continuous output space (lfr_plan_create)."""

    def __init__(self, space: Space, colouring: Colouring, kind: int):
        self.space, self.kind = space, kind
        # held so the colouring's id (the Space's plan cache key) stays unique
        self.colouring = colouring
        cic = np.asfortranarray(colouring.cells_in_colour.astype(np.int32))
        npc = np.ascontiguousarray(colouring.ncells_per_colour, dtype=np.int32)
        self._keep = (cic, npc)
        h = _P()
        check(lib().lfr_plan_create(space.ctx.h, space.h, cic.ctypes.data_as(_P),
                                    npc.ctypes.data_as(_P), colouring.ncolours,
                                    cic.shape[1], kind, C.byref(h)), space.ctx.h)
        self.h = h

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int64()
        check(lib().lfr_plan_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"ncolumns": a.value, "nclasses": b.value, "nbytes": c.value}

    def close(self):
        if getattr(self, "h", None):
            lib().lfr_plan_destroy(self.h)
            self.h = None


@dataclass
class LMAOperator:
    """operator_type: local_stencil(ncell_3d, ndf_to, ndf_from) between two spaces."""

    matrix: torch.Tensor          # (ndf_from, ndf_to, ncell_3d), contiguous
    to_space: Space
    from_space: Space
    colouring: Colouring | None = None
    apply_bc: bool = False        # solvers: enforce_bc after every apply (mass_matrix_operator apply_bc)
    packed: "PackedOperator | None" = None  # solvers: apply through this kind-0 pack
    _cache: dict = field(default_factory=dict)

    def __post_init__(self):
        nt, nf = self.to_space.ndf, self.from_space.ndf
        n3 = self.to_space.fs.ncell_3d
        if tuple(self.matrix.shape) != (nf, nt, n3):
            raise LfrError(1, f"operator matrix shape {tuple(self.matrix.shape)} != {(nf, nt, n3)}")
        if not self.matrix.is_contiguous():
            raise LfrError(1, "operator matrix must be contiguous")
        if self.matrix.device != self.to_space.ctx.device:
            raise LfrError(1, "operator matrix on the wrong device")

    @property
    def ctx(self) -> Context:
        return self.to_space.ctx

    def with_matrix(self, matrix: torch.Tensor) -> "LMAOperator":
        return LMAOperator(matrix, self.to_space, self.from_space, self.colouring)


def _field(t: torch.Tensor, space: Space, name: str, dtype=None) -> None:
    if not isinstance(t, torch.Tensor):
        raise LfrError(1, f"{name} must be a torch tensor")
    if t.device != space.ctx.device:
        raise LfrError(1, f"{name} is on {t.device}, context is on {space.ctx.device}")
    if t.dim() != 1 or t.numel() != space.undf:
        raise LfrError(1, f"{name} has shape {tuple(t.shape)}, space has undf={space.undf}")
    if not t.is_contiguous():
        raise LfrError(1, f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise LfrError(1, f"{name} dtype {t.dtype} differs from the operator's {dtype}")


def matrix_vector(lhs: torch.Tensor, x: torch.Tensor, op: LMAOperator, overwrite: bool = False):
    """lhs += M x (GH_INC, colour order); overwrite=True gives lhs = M x."""
    if op.colouring is None:
        raise LfrError(1, "matrix_vector needs the operator's colouring")
    _field(lhs, op.to_space, "lhs", op.matrix.dtype)
    _field(x, op.from_space, "x", op.matrix.dtype)
    plan = op.to_space.plan(op.colouring, 0)
    ctx = op.ctx
    ctx.sync_stream()
    check(lib().lfr_matrix_vector(ctx.h, _prec(lhs), _P(lhs.data_ptr()), _P(x.data_ptr()),
                                  _P(op.matrix.data_ptr()), op.to_space.h, op.from_space.h,
                                  plan.h, int(overwrite)), ctx.h)
    return lhs


def _ndata_field(t: torch.Tensor, space: Space, name: str, dtype) -> int:
    if not isinstance(t, torch.Tensor) or t.device != space.ctx.device:
        raise LfrError(1, f"{name} must be a tensor on {space.ctx.device}")
    if t.dim() != 2 or t.shape[0] != space.undf or not t.is_contiguous():
        raise LfrError(1, f"{name} must be a contiguous (undf={space.undf}, ndata) tensor, got "
                          f"{tuple(t.shape)}")
    if t.dtype != dtype:
        raise LfrError(1, f"{name} dtype {t.dtype} differs from the operator's {dtype}")
    return t.shape[1]


def matrix_vector_ndata(lhs: torch.Tensor, x: torch.Tensor, op: LMAOperator, overwrite: bool = False,
                        transpose: bool = False, colouring: Colouring | None = None):
    """matrix_vector (or, with transpose=True, transpose_matrix_vector) on a
    multidata field: lhs[:, r] (+)= A x[:, r] for every column r of the
    (undf, ndata) tensors ("ndata first" storage); each field bit-identical to
    the single-field call."""
    col = colouring or op.colouring
    if col is None:
        raise LfrError(1, "matrix_vector_ndata needs a colouring of the output space")
    lsp, xsp = (op.from_space, op.to_space) if transpose else (op.to_space, op.from_space)
    nd = _ndata_field(lhs, lsp, "lhs", op.matrix.dtype)
    if _ndata_field(x, xsp, "x", op.matrix.dtype) != nd:
        raise LfrError(1, "lhs and x have different ndata")
    plan = lsp.plan(col, 1 if transpose else 0)
    ctx = op.ctx
    ctx.sync_stream()
    check(lib().lfr_matrix_vector_ndata(ctx.h, _prec(lhs), _P(lhs.data_ptr()), _P(x.data_ptr()),
                                        _P(op.matrix.data_ptr()), lsp.h, xsp.h, plan.h, nd,
                                        int(overwrite)), ctx.h)
    return lhs


def dg_matrix_vector(lhs: torch.Tensor, x: torch.Tensor, op: LMAOperator):
    """lhs = M x on a discontinuous output space."""
    _field(lhs, op.to_space, "lhs", op.matrix.dtype)
    _field(x, op.from_space, "x", op.matrix.dtype)
    ctx = op.ctx
    ctx.sync_stream()
    check(lib().lfr_dg_matrix_vector(ctx.h, _prec(lhs), _P(lhs.data_ptr()), _P(x.data_ptr()),
                                     _P(op.matrix.data_ptr()), op.to_space.h, op.from_space.h),
          ctx.h)
    return lhs


def dg_inc_matrix_vector(lhs: torch.Tensor, x: torch.Tensor, op: LMAOperator):
    """lhs += M x on a discontinuous output space."""
    _field(lhs, op.to_space, "lhs", op.matrix.dtype)
    _field(x, op.from_space, "x", op.matrix.dtype)
    ctx = op.ctx
    ctx.sync_stream()
    check(lib().lfr_dg_inc_matrix_vector(ctx.h, _prec(lhs), _P(lhs.data_ptr()),
                                         _P(x.data_ptr()), _P(op.matrix.data_ptr()),
                                         op.to_space.h, op.from_space.h), ctx.h)
    return lhs


def transpose_matrix_vector(lhs: torch.Tensor, x: torch.Tensor, op: LMAOperator,
                            colouring: Colouring | None = None, overwrite: bool = False):
    """lhs += M^T x with op.matrix = local_stencil(ncell_3d, ndf_x, ndf_lhs).

    ``op.to_space`` is x's space and ``op.from_space`` is lhs' space, so the
    same LMAOperator that maps from_space -> to_space in ``matrix_vector``
    is applied transposed here (the adjoint).
    """
    col = colouring or op.colouring
    if col is None:
        raise LfrError(1, "transpose_matrix_vector needs a colouring of lhs' space")
    lsp, xsp = op.from_space, op.to_space
    _field(lhs, lsp, "lhs", op.matrix.dtype)
    _field(x, xsp, "x", op.matrix.dtype)
    plan = lsp.plan(col, 1)
    ctx = op.ctx
    ctx.sync_stream()
    check(lib().lfr_transpose_matrix_vector(ctx.h, _prec(lhs), _P(lhs.data_ptr()),
                                            _P(x.data_ptr()), _P(op.matrix.data_ptr()),
                                            lsp.h, xsp.h, plan.h, int(overwrite)), ctx.h)
    return lhs


class PackedOperator:
    """Gather-order copy of an LMAOperator for repeated applies (lfr_lma_pack).

    kind 0 packs the operator of ``matrix_vector`` (output on ``to_space``),
    kind 1 that of ``transpose_matrix_vector`` (output on ``from_space``).
    ``apply`` gives the same bits as the unpacked call; the pack holds its own
    copy of the matrix (``nbytes``), so ``op.matrix`` may change or go away
    afterwards without affecting it.
    """

    def __init__(self, op: LMAOperator, kind: int = 0, colouring: Colouring | None = None):
        col = colouring or op.colouring
        if col is None:
            raise LfrError(1, "packing needs a colouring of the output space")
        if kind not in (0, 1):
            raise LfrError(1, f"kind must be 0 or 1, got {kind}")
        self.kind = kind
        self.out_space, self.x_space = ((op.to_space, op.from_space) if kind == 0
                                        else (op.from_space, op.to_space))
        self.plan = self.out_space.plan(col, kind)
        self.dtype = op.matrix.dtype
        self.ctx = op.ctx
        self.ctx.sync_stream()
        h = _P()
        check(lib().lfr_lma_pack(self.ctx.h, _prec(op.matrix), _P(op.matrix.data_ptr()),
                                 self.x_space.h, self.plan.h, C.byref(h)), self.ctx.h)
        self.h = h

    @property
    def nbytes(self) -> int:
        n = C.c_longlong()
        check(lib().lfr_lma_packed_info(self.h, C.byref(n)))
        return n.value

    def apply(self, lhs: torch.Tensor, x: torch.Tensor, overwrite: bool = False) -> torch.Tensor:
        """lhs (+)= A x (kind 0) or A^T x (kind 1), bit-identical to the unpacked apply."""
        _field(lhs, self.out_space, "lhs", self.dtype)
        _field(x, self.x_space, "x", self.dtype)
        self.ctx.sync_stream()
        check(lib().lfr_matrix_vector_packed(self.ctx.h, self.h, _P(lhs.data_ptr()),
                                             _P(x.data_ptr()), int(overwrite)), self.ctx.h)
        return lhs

    def close(self):
        if getattr(self, "h", None):
            lib().lfr_lma_packed_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- vector algebra ---------------------------------------------------------

def enforce_bc(field_: torch.Tensor, space: Space) -> torch.Tensor:
    """invoke(enforce_bc_kernel_type(field)): zero the bottom/top boundary dofs."""
    _field(field_, space, "field")
    ctx = space.ctx
    ctx.sync_stream()
    check(lib().lfr_enforce_bc(ctx.h, _prec(field_), _P(field_.data_ptr()), space.h), ctx.h)
    return field_


def _vec(ctx: Context, *ts: torch.Tensor):
    n = ts[0].numel()
    for t in ts:
        if t.device != ctx.device or not t.is_contiguous() or t.numel() != n or t.dtype != ts[0].dtype:
            raise LfrError(1, "vector operands must be contiguous, same length/dtype/device")
    ctx.sync_stream()
    return _prec(ts[0]), n


def setval(ctx, x, c):
    p, n = _vec(ctx, x)
    check(lib().lfr_vec_setval(ctx.h, p, n, _P(x.data_ptr()), float(c)), ctx.h)


def copy(ctx, y, x):
    p, n = _vec(ctx, y, x)
    check(lib().lfr_vec_copy(ctx.h, p, n, _P(y.data_ptr()), _P(x.data_ptr())), ctx.h)


def axpy(ctx, a, x, y):
    """y = y + a x"""
    p, n = _vec(ctx, x, y)
    check(lib().lfr_vec_axpy(ctx.h, p, n, float(a), _P(x.data_ptr()), _P(y.data_ptr())), ctx.h)


def aypx(ctx, a, x, y):
    """y = a y + x"""
    p, n = _vec(ctx, x, y)
    check(lib().lfr_vec_aypx(ctx.h, p, n, float(a), _P(x.data_ptr()), _P(y.data_ptr())), ctx.h)


def axpby(ctx, a, x, b, y, z):
    """z = a x + b y"""
    p, n = _vec(ctx, x, y, z)
    check(lib().lfr_vec_axpby(ctx.h, p, n, float(a), _P(x.data_ptr()), float(b),
                              _P(y.data_ptr()), _P(z.data_ptr())), ctx.h)


def scale(ctx, a, x):
    p, n = _vec(ctx, x)
    check(lib().lfr_vec_scale(ctx.h, p, n, float(a), _P(x.data_ptr())), ctx.h)


def divide(ctx, z, x, y):
    p, n = _vec(ctx, z, x, y)
    check(lib().lfr_vec_divide(ctx.h, p, n, _P(z.data_ptr()), _P(x.data_ptr()), _P(y.data_ptr())), ctx.h)


def multiply(ctx, z, x, y):
    p, n = _vec(ctx, z, x, y)
    check(lib().lfr_vec_multiply(ctx.h, p, n, _P(z.data_ptr()), _P(x.data_ptr()), _P(y.data_ptr())), ctx.h)


def convert(ctx, y, x):
    if x.numel() != y.numel() or x.device != ctx.device or y.device != ctx.device:
        raise LfrError(1, "convert: operands must match in length and device")
    ctx.sync_stream()
    check(lib().lfr_vec_convert(ctx.h, _prec(y), _P(y.data_ptr()), _prec(x), _P(x.data_ptr()),
                                x.numel()), ctx.h)


def dot(ctx, x, y) -> float:
    """<x, y> accumulated in FP64 (deterministic)."""
    p, n = _vec(ctx, x, y)
    out = C.c_double()
    check(lib().lfr_vec_dot(ctx.h, p, n, _P(x.data_ptr()), _P(y.data_ptr()), C.byref(out)), ctx.h)
    return out.value


# --- column-wise operators --------------------------------------------------

class MultiOperator:
    """An LMA operator prepared for 16-field applies on MFMA (lfr_multi_create).

    mprec 4: FP32 matrix and FP32 MFMA; mprec 2: BF16 matrix and inputs, FP32
    accumulation.  Fields are (undf, 16) FP32 tensors (the 16 values of a dof
    contiguous)."""

    NRHS = 16

    def __init__(self, op: LMAOperator, mprec: int = 4):
        self.op, self.mprec = op, mprec
        ctx = op.ctx
        col = op.colouring
        if col is not None:
            cic = np.asfortranarray(col.cells_in_colour.astype(np.int32))
            npc = np.ascontiguousarray(col.ncells_per_colour, dtype=np.int32)
            args = (cic.ctypes.data_as(_P), npc.ctypes.data_as(_P), col.ncolours, cic.shape[1])
            self._keep = (cic, npc)
        else:
            args = (None, None, 0, 0)
        h = _P()
        check(lib().lfr_multi_create(ctx.h, mprec, _prec(op.matrix), _P(op.matrix.data_ptr()),
                                     op.to_space.h, op.from_space.h, *args, C.byref(h)), ctx.h)
        self.h = h

    def info(self):
        nb, s, p, nlch = C.c_int64(), C.c_int(), C.c_int(), C.c_int()
        check(lib().lfr_multi_info(self.h, C.byref(nb), C.byref(s), C.byref(p), C.byref(nlch)))
        return {"packed_bytes": nb.value, "slots": s.value, "vertical_pairs": p.value,
                "launches": nlch.value}

    def _check(self, t: torch.Tensor, space: Space, name: str):
        shape = (space.undf, self.NRHS)
        if tuple(t.shape) != shape or t.dtype != torch.float32 or not t.is_contiguous() \
                or t.device != space.ctx.device:
            raise LfrError(1, f"{name} must be a contiguous float32 tensor of shape {shape}")

    def apply(self, lhs: torch.Tensor, x: torch.Tensor, overwrite: bool = True):
        """lhs (+)= A x for the 16 fields."""
        self._check(lhs, self.op.to_space, "lhs")
        self._check(x, self.op.from_space, "x")
        ctx = self.op.ctx
        ctx.sync_stream()
        check(lib().lfr_matrix_vector_multi(ctx.h, self.h, _P(lhs.data_ptr()), _P(x.data_ptr()),
                                            self.NRHS, int(overwrite)), ctx.h)

    def close(self):
        if getattr(self, "h", None):
            lib().lfr_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CMAOperator:
    """columnwise_operator_type on one space: columnwise_matrix(bw, nrow, ncell_2d)
    as a tensor of shape (ncell_2d, nrow, bw) plus indirection_dofmap(nrow)."""

    def __init__(self, matrix: torch.Tensor, space: Space, indirection: np.ndarray):
        nc = space.fs.ncells
        if matrix.dim() != 3 or matrix.shape[0] != nc:
            raise LfrError(1, f"CMA matrix shape {tuple(matrix.shape)} does not match ncell={nc}")
        if not matrix.is_contiguous() or matrix.device != space.ctx.device:
            raise LfrError(1, "CMA matrix must be contiguous and on the context's device")
        self.matrix, self.space = matrix, space
        self.nrow, self.bandwidth = int(matrix.shape[1]), int(matrix.shape[2])
        ind = np.ascontiguousarray(indirection, dtype=np.int32)
        if ind.shape != (self.nrow,):
            raise LfrError(1, "indirection_dofmap must have nrow entries")
        base = space.fs.dofmap[:, 0].astype(np.int64)
        if ind.min() < 1 or base.min() + ind.max() - 1 > space.undf:
            raise LfrError(1, "indirection_dofmap addresses dofs outside the space")
        self.indirection = torch.from_numpy(ind).to(space.ctx.device)


def columnwise_op_appinv(lhs: torch.Tensor, x: torch.Tensor, op: CMAOperator,
                         fp64_pivots: bool = False):
    """lhs = A^{-1} x column by column (tridiagonal Thomas solve; overwrites lhs)."""
    _field(lhs, op.space, "lhs", op.matrix.dtype)
    _field(x, op.space, "x", op.matrix.dtype)
    ctx = op.space.ctx
    ctx.sync_stream()
    check(lib().lfr_columnwise_op_appinv(ctx.h, _prec(lhs), _P(lhs.data_ptr()), _P(x.data_ptr()),
                                         _P(op.matrix.data_ptr()), op.bandwidth, op.nrow,
                                         op.space.fs.ncells, op.space.h,
                                         _P(op.indirection.data_ptr()), int(fp64_pivots)), ctx.h)
    return lhs


def tri_solve(y: torch.Tensor, x: torch.Tensor, tri_0: torch.Tensor, tri_plus: torch.Tensor,
              tri_minus: torch.Tensor, space: Space) -> torch.Tensor:
    """invoke(tri_solve_kernel_type(y, x, tri_0, tri_plus, tri_minus)): the
    vertical tridiagonal solve on every column of a lowest-order W3 space;
    y is overwritten."""
    dt = x.dtype
    for name, t in (("y", y), ("x", x), ("tri_0", tri_0), ("tri_plus", tri_plus), ("tri_minus", tri_minus)):
        _field(t, space, name, dt)
    ctx = space.ctx
    ctx.sync_stream()
    check(lib().lfr_tri_solve(ctx.h, _prec(y), _P(y.data_ptr()), _P(x.data_ptr()), _P(tri_0.data_ptr()),
                              _P(tri_plus.data_ptr()), _P(tri_minus.data_ptr()), space.h), ctx.h)
    return y


def operator_tri_solve(y: torch.Tensor, x: torch.Tensor, matrix: torch.Tensor, space: Space) -> torch.Tensor:
    """invoke(operator_tri_solve_kernel_type(y, x, matrix)): the column
    tridiagonal of a lowest-order Wtheta/W2v space taken from its 2x2 LMA
    operator (matrix (2, 2, ncell_3d), the local_stencil memory), then the
    Thomas solve; y is overwritten."""
    _field(y, space, "y", x.dtype)
    _field(x, space, "x", x.dtype)
    n3 = space.fs.mesh.ncells * space.fs.nlayers
    if tuple(matrix.shape) != (2, 2, n3) or not matrix.is_contiguous() or matrix.dtype != x.dtype:
        raise LfrError(1, f"matrix must be a contiguous (2, 2, {n3}) {x.dtype} tensor")
    ctx = space.ctx
    ctx.sync_stream()
    check(lib().lfr_operator_tri_solve(ctx.h, _prec(y), _P(y.data_ptr()), _P(x.data_ptr()),
                                       _P(matrix.data_ptr()), space.h), ctx.h)
    return y


def compute_mass_matrix_w0(chi: torch.Tensor, space: Space, nq: int = 3, stiffness: bool = False):
    """compute_mass_matrix_w_scalar for lowest-order W0 on the device: chi is
    the (3, undf) FP64 coordinate field (chi1, chi2, chi3) on ``space``;
    returns the local_stencil (8, 8, ncell_3d) mass matrix (and the stiffness
    matrix with ``stiffness=True``)."""
    if chi.dim() != 2 or chi.shape[0] != 3 or chi.shape[1] != space.undf or chi.dtype != torch.float64 \
            or not chi.is_contiguous() or chi.device != space.ctx.device:
        raise LfrError(1, f"chi must be a contiguous (3, {space.undf}) float64 tensor on {space.ctx.device}")
    n3 = space.fs.mesh.ncells * space.fs.nlayers
    mass = torch.empty((8, 8, n3), dtype=torch.float64, device=chi.device)
    stiff = torch.empty_like(mass) if stiffness else None
    ctx = space.ctx
    ctx.sync_stream()
    check(lib().lfr_compute_mass_matrix_w0(ctx.h, int(nq), _P(chi[0].data_ptr()), _P(chi[1].data_ptr()),
                                           _P(chi[2].data_ptr()), space.h, _P(mass.data_ptr()),
                                           _P(stiff.data_ptr()) if stiff is not None else None), ctx.h)
    return (mass, stiff) if stiffness else mass


def compute_mass_matrix(kind: int, basis, diff_basis_chi, wqp_h, wqp_v, chi: torch.Tensor, chi_space: Space):
    """Mass matrix of a lowest-order space on the device from basis tables
    at the quadrature points: kind 0 = scalar space
    (compute_mass_matrix_w_scalar, basis (1, ndf, nqp_h, nqp_v)), kind 1 =
    W2 / W2H / W2V with the Piola map (compute_mass_matrix_w2, basis
    (3, ndf, nqp_h, nqp_v)); diff_basis_chi (3, 8, nqp_h, nqp_v) of the W0
    chi space; chi the (3, undf) FP64 coordinate field on ``chi_space``.
    Returns the local_stencil (ndf, ndf, ncell_3d) in FP64."""
    basis = np.asarray(basis, np.float64)
    dchi = np.asarray(diff_basis_chi, np.float64)
    wh = np.ascontiguousarray(wqp_h, np.float64)
    wv = np.ascontiguousarray(wqp_v, np.float64)
    if basis.ndim != 4 or dchi.shape != (3, 8) + basis.shape[2:] or wh.shape != basis.shape[2:3] \
            or wv.shape != basis.shape[3:4]:
        raise LfrError(1, "basis (ncomp, ndf, nqp_h, nqp_v), diff_basis_chi (3, 8, nqp_h, nqp_v), "
                          "wqp_h (nqp_h,), wqp_v (nqp_v,) expected")
    if chi.dim() != 2 or chi.shape[0] != 3 or chi.shape[1] != chi_space.undf or chi.dtype != torch.float64 \
            or not chi.is_contiguous() or chi.device != chi_space.ctx.device:
        raise LfrError(1, f"chi must be a contiguous (3, {chi_space.undf}) float64 tensor on {chi_space.ctx.device}")
    ncomp, ndf, nh, nv = basis.shape
    if ncomp != (1 if kind == 0 else 3):
        raise LfrError(1, f"kind {kind} needs a basis with {1 if kind == 0 else 3} components, got {ncomp}")
    b = np.ascontiguousarray(basis.transpose(3, 2, 1, 0))      # C order of the Fortran (c, df, h, v)
    d = np.ascontiguousarray(dchi.transpose(3, 2, 1, 0))
    n3 = chi_space.fs.mesh.ncells * chi_space.fs.nlayers
    mass = torch.empty((ndf, ndf, n3), dtype=torch.float64, device=chi.device)
    ctx = chi_space.ctx
    ctx.sync_stream()
    dp = lambda a: a.ctypes.data_as(_P)
    check(lib().lfr_compute_mass_matrix(ctx.h, int(kind), int(ndf), int(nh), int(nv), dp(b), dp(d), dp(wh),
                                        dp(wv), _P(chi[0].data_ptr()), _P(chi[1].data_ptr()),
                                        _P(chi[2].data_ptr()), chi_space.h, _P(mass.data_ptr())), ctx.h)
    return mass


class Intergrid:
    """The intergrid map of a coarse and a fine lowest-order space (W3,
    Wtheta or W2; lfr_intergrid_create): cell_map (ncell_coarse, nfy, nfx),
    1-based fine cells, as the reference's mesh-to-mesh connectivity."""

    def __init__(self, coarse_space: Space, fine_space: Space, cell_map):
        cm = np.ascontiguousarray(cell_map, dtype=np.int32)
        if cm.ndim != 3 or cm.shape[0] != coarse_space.fs.mesh.ncells:
            raise LfrError(1, f"cell_map must be (ncell_coarse, nfy, nfx), got {cm.shape}")
        self.coarse_space, self.fine_space = coarse_space, fine_space
        h = _P()
        ctx = coarse_space.ctx
        check(lib().lfr_intergrid_create(ctx.h, coarse_space.h, fine_space.h, cm.ctypes.data_as(_P),
                                         int(cm.shape[2]), int(cm.shape[1]), C.byref(h)), ctx.h)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lfr_intergrid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _transfer(name, out, inp, ig: Intergrid):
    if out.dtype != inp.dtype:
        raise LfrError(1, "the two fields must have one dtype")
    fo, fi = (ig.coarse_space, ig.fine_space) if name == "restrict" else (ig.fine_space, ig.coarse_space)
    _field(out, fo, "output", inp.dtype)
    _field(inp, fi, "input", inp.dtype)
    ctx = ig.coarse_space.ctx
    ctx.sync_stream()
    check(getattr(lib(), f"lfr_{name}_scalar")(ctx.h, _prec(out), _P(out.data_ptr()), _P(inp.data_ptr()), ig.h),
          ctx.h)
    return out


def restrict_scalar(coarse: torch.Tensor, fine: torch.Tensor, ig: Intergrid):
    """restrict_scalar_unweighted over every coarse cell (W3 or Wtheta, lowest
    order): coarse = the mean of the children."""
    return _transfer("restrict", coarse, fine, ig)


def prolong_scalar(fine: torch.Tensor, coarse: torch.Tensor, ig: Intergrid):
    """prolong_scalar_unweighted: fine += coarse on every child."""
    return _transfer("prolong", fine, coarse, ig)


def restrict_w2(coarse: torch.Tensor, fine: torch.Tensor, ig: Intergrid):
    """restrict_w2 over an intergrid map between lowest-order W2 spaces: each
    coarse face receives the sum of the fine faces on it."""
    if coarse.dtype != fine.dtype:
        raise LfrError(1, "the two fields must have one dtype")
    _field(coarse, ig.coarse_space, "coarse", fine.dtype)
    _field(fine, ig.fine_space, "fine", fine.dtype)
    ctx = ig.coarse_space.ctx
    ctx.sync_stream()
    check(lib().lfr_restrict_w2(ctx.h, _prec(coarse), _P(coarse.data_ptr()), _P(fine.data_ptr()), ig.h), ctx.h)
    return coarse


def prolong_w2(fine: torch.Tensor, coarse: torch.Tensor, ig: Intergrid):
    """prolong_w2 with its weights (square children): every fine face of the
    coarse cells is overwritten with its share of the coarse fluxes."""
    if coarse.dtype != fine.dtype:
        raise LfrError(1, "the two fields must have one dtype")
    _field(fine, ig.fine_space, "fine", coarse.dtype)
    _field(coarse, ig.coarse_space, "coarse", coarse.dtype)
    ctx = ig.coarse_space.ctx
    ctx.sync_stream()
    check(lib().lfr_prolong_w2(ctx.h, _prec(fine), _P(fine.data_ptr()), _P(coarse.data_ptr()), ig.h), ctx.h)
    return fine


_DERHAM_TABLES = ("diff_basis_chi", "basis_w0", "diff_basis_w0", "basis_w1", "diff_basis_w1", "basis_w2",
                  "diff_basis_w2", "basis_w2b", "diff_basis_w2b", "basis_w3", "basis_wt")
# (output, ndf_to, ndf_from, tables it needs)
DERHAM_OUTPUTS = (("mm0", 8, 8, ("basis_w0",)), ("mm1", 12, 12, ("basis_w1",)), ("mm2", 6, 6, ("basis_w2",)),
                  ("mm2b", 6, 6, ("basis_w2b",)), ("mm3", 1, 1, ("basis_w3",)), ("mmt", 2, 2, ("basis_wt",)),
                  ("grad", 12, 8, ("basis_w1", "diff_basis_w0")), ("curl", 6, 12, ("basis_w2", "diff_basis_w1")),
                  ("div", 1, 6, ("basis_w3", "diff_basis_w2")),
                  ("broken_div", 1, 6, ("basis_w3", "diff_basis_w2b")))


class _DerhamTables(C.Structure):
    _fields_ = [("nqp_h", C.c_int), ("nqp_v", C.c_int), ("wqp_h", _P), ("wqp_v", _P)] + \
               [(n, _P) for n in _DERHAM_TABLES]


def compute_derham_matrices(tables: dict, chi: torch.Tensor, chi_space: Space, outputs=None):
    """compute_derham_matrices on the device, lowest order: ``tables`` holds
    wqp_h, wqp_v and the basis tables named like the kernel's arguments
    (basis_w0 (1, 8, h, v), diff_basis_w0 (3, 8, h, v), basis_w1 / diff_basis_w1
    (3, 12, h, v), basis_w2 / basis_w2b (3, 6, h, v), diff_basis_w2 /
    diff_basis_w2b (1, 6, h, v), basis_w3 (1, 1, h, v), basis_wt (1, 2, h, v),
    diff_basis_chi (3, 8, h, v)); ``outputs`` names which of mm0, mm1, mm2,
    mm2b, mm3, mmt, grad, curl, div, broken_div to compute (all by default).
    Returns {name: local_stencil (ndf_from, ndf_to, ncell_3d) FP64 tensor}."""
    names = [o[0] for o in DERHAM_OUTPUTS] if outputs is None else list(outputs)
    shapes = {o[0]: o[1:] for o in DERHAM_OUTPUTS}
    bad = [n for n in names if n not in shapes]
    if bad:
        raise LfrError(1, f"unknown de Rham outputs {bad}")
    if chi.dim() != 2 or chi.shape[0] != 3 or chi.shape[1] != chi_space.undf or chi.dtype != torch.float64 \
            or not chi.is_contiguous() or chi.device != chi_space.ctx.device:
        raise LfrError(1, f"chi must be a contiguous (3, {chi_space.undf}) float64 tensor on {chi_space.ctx.device}")
    wh = np.ascontiguousarray(tables["wqp_h"], np.float64)
    wv = np.ascontiguousarray(tables["wqp_v"], np.float64)
    nh, nv = wh.size, wv.size
    want = {"diff_basis_chi": (3, 8)}
    for n in names:
        for tname in shapes[n][2]:
            nd = {"w0": 8, "w1": 12, "w2": 6, "w2b": 6, "w3": 1, "wt": 2}[tname.rsplit("_", 1)[1]]
            nc = 1 if tname in ("basis_w0", "basis_w3", "basis_wt", "diff_basis_w2", "diff_basis_w2b") else 3
            want[tname] = (nc, nd)
    keep, st = [], _DerhamTables(nqp_h=nh, nqp_v=nv)
    for tname, lead in want.items():
        a = np.asarray(tables[tname], np.float64)
        if a.shape != lead + (nh, nv):
            raise LfrError(1, f"{tname} must have shape {lead + (nh, nv)}, got {a.shape}")
        a = np.ascontiguousarray(a.transpose(3, 2, 1, 0))          # C order of the Fortran table
        keep.append(a)
        setattr(st, tname, a.ctypes.data_as(_P))
    st.wqp_h, st.wqp_v = wh.ctypes.data_as(_P), wv.ctypes.data_as(_P)
    n3 = chi_space.fs.mesh.ncells * chi_space.fs.nlayers
    out = {n: torch.empty((shapes[n][1], shapes[n][0], n3), dtype=torch.float64, device=chi.device)
           for n in names}
    ptr = [_P(out[o[0]].data_ptr()) if o[0] in out else None for o in DERHAM_OUTPUTS]
    ctx = chi_space.ctx
    ctx.sync_stream()
    check(lib().lfr_compute_derham_matrices(ctx.h, C.byref(st), _P(chi[0].data_ptr()), _P(chi[1].data_ptr()),
                                            _P(chi[2].data_ptr()), chi_space.h, *ptr), ctx.h)
    return out


class ColumnwiseGeometry:
    """Banded column-operator geometry between two horizontally discontinuous
    spaces (lfr_cma_create), from lfric_core_amd.cma.cma_geometry."""

    def __init__(self, to_space: Space, from_space: Space, geom=None):
        from .cma import cma_geometry
        g = geom if geom is not None else cma_geometry(to_space.fs, from_space.fs)
        self.geom, self.to_space, self.from_space = g, to_space, from_space
        bt = np.ascontiguousarray(g.banded_to, dtype=np.int32)       # (nl, ndf) = Fortran (ndf, nl)
        bf = np.ascontiguousarray(g.banded_from, dtype=np.int32)
        it = np.ascontiguousarray(g.indirection_to, dtype=np.int32)
        if_ = np.ascontiguousarray(g.indirection_from, dtype=np.int32)
        h = _P()
        ctx = to_space.ctx
        check(lib().lfr_cma_create(ctx.h, to_space.fs.nlayers, g.nrow, g.ncol, g.bandwidth, g.alpha,
                                   g.beta, g.gamma_m, g.gamma_p, to_space.ndf, bt.ctypes.data_as(_P),
                                   it.ctypes.data_as(_P), from_space.ndf, bf.ctypes.data_as(_P),
                                   if_.ctypes.data_as(_P), C.byref(h)), ctx.h)
        self.h = h
        self.indirection_to = torch.from_numpy(it).to(ctx.device)

    def empty_matrix(self, dtype=torch.float32) -> torch.Tensor:
        g = self.geom
        return torch.empty((self.to_space.fs.ncells, g.nrow, g.bandwidth), dtype=dtype,
                           device=self.to_space.ctx.device)

    def close(self):
        if getattr(self, "h", None):
            lib().lfr_cma_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def columnwise_op_asm(cm: torch.Tensor, lma: LMAOperator, geom: ColumnwiseGeometry) -> torch.Tensor:
    """columnwise_matrix(bw, nrow, ncell_2d) (tensor (ncell_2d, nrow, bw)) assembled from an LMA operator."""
    g = geom.geom
    shape = (geom.to_space.fs.ncells, g.nrow, g.bandwidth)
    if tuple(cm.shape) != shape or cm.dtype != lma.matrix.dtype or not cm.is_contiguous():
        raise LfrError(1, f"cm must be a contiguous {lma.matrix.dtype} tensor of shape {shape}")
    ctx = geom.to_space.ctx
    ctx.sync_stream()
    check(lib().lfr_columnwise_op_asm(ctx.h, _prec(cm), geom.h, _P(lma.matrix.data_ptr()),
                                      _P(cm.data_ptr()), geom.to_space.fs.ncells), ctx.h)
    return cm


def columnwise_op_app(lhs: torch.Tensor, x: torch.Tensor, cm: torch.Tensor, geom: ColumnwiseGeometry):
    """lhs += A x for a banded column operator."""
    _field(lhs, geom.to_space, "lhs", cm.dtype)
    _field(x, geom.from_space, "x", cm.dtype)
    ctx = geom.to_space.ctx
    ctx.sync_stream()
    check(lib().lfr_columnwise_op_app(ctx.h, _prec(cm), geom.h, _P(lhs.data_ptr()), _P(x.data_ptr()),
                                      _P(cm.data_ptr()), geom.to_space.h, geom.from_space.h), ctx.h)
    return lhs
