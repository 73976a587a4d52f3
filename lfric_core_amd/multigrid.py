"""Multigrid hierarchy and V-cycle for the W3 Helmholtz-type column operator.

The operator on level l of a hierarchy of nested cubed spheres is
``A_l = T_l + sigma_l * L_h``: ``T_l`` a column tridiagonal (the CMA operator
that ``columnwise_op_appinv`` inverts,
components/science/source/kernel/columnwise_operators/sci_columnwise_op_appinv_kernel_mod.F90)
and ``L_h`` the horizontal graph Laplacian over the four face neighbours
(4 u - sum of neighbours).  The finest level draws (seed order: lambda for
every cell and layer, cell-major; then the sub-diagonal noise; then the
super-diagonal noise, all from one SplitMix64 stream)::

    diag  = 2 + h^2 * lambda,  lambda ~ U[0.1, 10],  h = 1 / nlayers
    sub   = -1 * (1 + 0.1 * U[-1, 1])   (0 in the bottom layer)
    super = -1 * (1 + 0.1 * U[-1, 1])   (0 in the top layer)

Coarse levels are Galerkin with the transfer pair of the reference
(cell-average restriction R, injection prolongation P): ``T_c`` is the
average of the children's tridiagonals (R T P, exactly) and ``sigma`` halves
per level (R L_h P = L_h_coarse / 2 on the cubed sphere).  The horizontal
term is this build's definition of the "Helmholtz-type" coupling: a purely
column-tridiagonal operator would be solved exactly by one line relaxation.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .mesh import cubed_sphere, intergrid_map
from .rng import Stream


@dataclass
class Level:
    edge_cells: int
    ncell: int
    cell_next: np.ndarray          # (ncell, 4) int32, 0-based
    children: np.ndarray | None    # (ncell, 4) int32, 0-based cells of the finer level (y outer, x inner)
    bands: np.ndarray              # (3, nl, ncell) float64: sub, diag, super, layer-major
    sigma: float


@dataclass
class Hierarchy:
    nlayers: int
    levels: list = field(default_factory=list)

    @property
    def fine_undf(self) -> int:
        return self.levels[0].ncell * self.nlayers


def build_hierarchy(edge_cells=(192, 96, 48, 24), nlayers: int = 70, seed: int = 4,
                    sigma0: float = 0.25) -> Hierarchy:
    nl = nlayers
    h = 1.0 / nl
    H = Hierarchy(nl)
    st = Stream(seed)
    for li, n in enumerate(edge_cells):
        mesh = cubed_sphere(n)
        nc = mesh.ncells
        cn = (mesh.cell_next - 1).astype(np.int32)
        if li == 0:
            lam = st.uniform(nc * nl, 0.1, 10.0).reshape(nc, nl)
            sub = -1.0 * (1.0 + 0.1 * st.uniform(nc * nl).reshape(nc, nl))
            sup = -1.0 * (1.0 + 0.1 * st.uniform(nc * nl).reshape(nc, nl))
            sub[:, 0] = 0.0
            sup[:, -1] = 0.0
            bands = np.stack([sub.T, (2.0 + h * h * lam).T, sup.T])    # (3, nl, nc)
            children = None
        else:
            prev = H.levels[-1]
            m = intergrid_map(n, edge_cells[li - 1]).reshape(nc, -1) - 1   # y outer, x inner
            children = np.ascontiguousarray(m, dtype=np.int32)
            fb = prev.bands
            acc = np.zeros((3, nl, nc))
            for q in range(m.shape[1]):
                acc = acc + fb[:, :, m[:, q]] * 0.25
            bands = acc
        sigma = sigma0 / (2.0 ** li)
        H.levels.append(Level(n, nc, np.ascontiguousarray(cn), children,
                              np.ascontiguousarray(bands), sigma))
    return H


class Multigrid:
    """Device V-cycle (lfr_mg_create / lfr_mg_solve) over nrhs right-hand sides."""

    def __init__(self, ctx, H: Hierarchy, prec: int = 4, nrhs: int = 1, omega: float = 0.8,
                 nu1: int = 2, nu2: int = 2, ncoarse: int = 8):
        from ._ffi import check, lib
        self.ctx, self.H, self.prec, self.nrhs = ctx, H, prec, nrhs
        L = len(H.levels)
        ncell = (C.c_int * L)(*[lv.ncell for lv in H.levels])
        self._keep = []
        cnp = (C.c_void_p * L)()
        chp = (C.c_void_p * L)()
        bp = (C.c_void_p * L)()
        for i, lv in enumerate(H.levels):
            cn = np.ascontiguousarray(lv.cell_next, dtype=np.int32)
            b = np.ascontiguousarray(lv.bands, dtype=np.float64)
            self._keep += [cn, b]
            cnp[i] = cn.ctypes.data
            bp[i] = b.ctypes.data
            if lv.children is not None:
                ch = np.ascontiguousarray(lv.children, dtype=np.int32)
                self._keep.append(ch)
                chp[i] = ch.ctypes.data
        sig = (C.c_double * L)(*[lv.sigma for lv in H.levels])
        h = C.c_void_p()
        check(lib().lfr_mg_create(ctx.h, prec, L, H.nlayers, nrhs, ncell, cnp, chp, bp, sig,
                                  float(omega), nu1, nu2, ncoarse, C.byref(h)), ctx.h)
        self.h = h
        bs = (C.c_int * L)()
        check(lib().lfr_mg_block_sizes(h, bs))
        self.block_sizes = list(bs)

    def solve(self, x, b, ncycles: int = 1, record: bool = True) -> np.ndarray | None:
        """x, b: (nrhs, undf) contiguous device tensors in LFRic W3 layout."""
        import torch
        from ._ffi import LfrError, check, lib
        dt = torch.float32 if self.prec == 4 else torch.float64
        shape = (self.nrhs, self.H.fine_undf)
        for t, nm in ((x, "x"), (b, "b")):
            if tuple(t.shape) != shape or t.dtype != dt or not t.is_contiguous() or t.device != self.ctx.device:
                raise LfrError(1, f"{nm} must be a contiguous {dt} tensor of shape {shape}")
        self.ctx.sync_stream()
        res = np.zeros((ncycles + 1, self.nrhs)) if record else None
        check(lib().lfr_mg_solve(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(b.data_ptr()),
                                 ncycles, res.ctypes.data_as(C.c_void_p) if record else None),
              self.ctx.h)
        return res

    def close(self):
        from ._ffi import lib
        if getattr(self, "h", None):
            lib().lfr_mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
