"""Columnwise (CMA) operator geometry for horizontally discontinuous spaces.

Restates columnwise_operator_type's set-up
(infrastructure/source/operator/columnwise_operator_mod.f90:168-191 initialiser,
:288-332 gcd, :350-396 build_dofmaps, :420-466 extract_mesh_fs_info):

* alpha = ni_from + nf_from, beta = ni_to + nf_to,
  gamma_m = (ndf_from - 1) * beta, gamma_p = (ndf_to - 1) * alpha, all four
  divided by gcd(alpha, beta) when it exceeds 1;
* bandwidth = 1 + ceiling(real(gamma_m + gamma_p) / real(beta)) in r_solver;
* nrow = nl*ni_to + (nl+1)*nf_to, ncol likewise for the from-space;
* column_banded_dofmap(df, k) = (k-1)*(ni + nf) + df and the indirection map
  indirection(banded(df, k)) = (dofmap(df) - dofmap(1) + 1) + (k - 1),
  taken from the first column (any column would do).

ni is the number of cell-interior dofs and nf the dofs per bottom/top face.
"""

from __future__ import annotations

from dataclasses import dataclass
from math import ceil, gcd

import numpy as np

from .function_space import FunctionSpace


@dataclass
class CMAGeometry:
    alpha: int
    beta: int
    gamma_m: int
    gamma_p: int
    bandwidth: int
    nrow: int
    ncol: int
    banded_to: np.ndarray       # column_banded_dofmap_to(ndf_to, nlayers), numpy (nlayers, ndf_to)
    banded_from: np.ndarray     # (nlayers, ndf_from)
    indirection_to: np.ndarray  # (nrow,) 1-based
    indirection_from: np.ndarray


def _nv(fs: FunctionSpace):
    ni, nf = fs.counts.vol, fs.counts.face_v
    if ni + 2 * nf != fs.ndf:
        raise ValueError(f"{fs.space} is not horizontally discontinuous")
    return ni, nf


def _maps(fs: FunctionSpace, ni: int, nf: int):
    nl = fs.nlayers
    m0 = fs.dofmap[0].astype(np.int64)
    if np.any(m0[1:] < m0[0]):
        raise ValueError("first dofmap entry is not the smallest")
    k = np.arange(1, nl + 1)[:, None]
    df = np.arange(1, fs.ndf + 1)[None, :]
    banded = (k - 1) * (ni + nf) + df                     # (nl, ndf)
    coldm = (m0[None, :] - m0[0] + 1) + (k - 1)
    nrow = nl * ni + (nl + 1) * nf
    ind = np.zeros(nrow, dtype=np.int32)
    ind[banded.ravel() - 1] = coldm.ravel()
    return banded.astype(np.int32), ind, nrow


def cma_geometry(fs_to: FunctionSpace, fs_from: FunctionSpace) -> CMAGeometry:
    ni_t, nf_t = _nv(fs_to)
    ni_f, nf_f = _nv(fs_from)
    alpha = ni_f + nf_f
    beta = ni_t + nf_t
    gamma_m = (fs_from.ndf - 1) * (ni_t + nf_t)
    gamma_p = (fs_to.ndf - 1) * (ni_f + nf_f)
    g = gcd(alpha, beta) if (alpha or beta) else 1
    if g > 1:
        alpha, beta, gamma_m, gamma_p = alpha // g, beta // g, gamma_m // g, gamma_p // g
    bw = 1 + int(ceil(np.float32(gamma_m + gamma_p) / np.float32(beta)))
    bt, it, nrow = _maps(fs_to, ni_t, nf_t)
    bf, if_, ncol = _maps(fs_from, ni_f, nf_f)
    return CMAGeometry(alpha, beta, gamma_m, gamma_p, bw, nrow, ncol, bt, bf, it, if_)
