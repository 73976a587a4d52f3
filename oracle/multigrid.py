"""V-cycle restated in numpy (TEST INFRASTRUCTURE ONLY), same operation order
as the device kernels: layer-major fields (nrhs, nl, ncell); residual
a x(k-1) + d x(k) + u x(k+1) + sigma (4 x - sum of the W,S,E,N neighbours);
line-Jacobi sweep with the Thomas recurrence of columnwise_op_appinv
(components/science/source/kernel/columnwise_operators/sci_columnwise_op_appinv_kernel_mod.F90:116-143)
on D = T + 4 sigma I; restriction as sci_restrict_scalar_unweighted_kernel_mod.F90
(children y outer, x inner, each times 1/4); injection prolongation as
sci_prolong_scalar_unweighted_kernel_mod.F90.

Pinned (tests/test_multigrid_oracle.py) by: Galerkin identity R A_f P = A_c
of the hierarchy, the Thomas solve against a dense solve, convergence of the
V-cycle iteration to the dense solution on a tiny hierarchy.
"""

from __future__ import annotations

import numpy as np


def apply_A(x, bands, cn, sigma):
    T = x.dtype.type
    a, d, u = (b.astype(x.dtype) for b in bands)
    xm = np.zeros_like(x)
    xp = np.zeros_like(x)
    xm[:, 1:] = x[:, :-1]
    xp[:, :-1] = x[:, 1:]
    ax = a * xm
    ax = ax + d * x
    ax = ax + u * xp
    s4 = x[..., cn[:, 0]] + x[..., cn[:, 1]]
    s5 = s4 + x[..., cn[:, 2]]
    sn = s5 + x[..., cn[:, 3]]
    hz = T(sigma) * (T(4) * x - sn)
    return ax + hz


def smooth(x, b, bands, cn, sigma, omega):
    T = x.dtype.type
    r = b - apply_A(x, bands, cn, sigma)
    a, d, u = (bb.astype(x.dtype) for bb in bands)
    dd = d + T(4) * T(sigma)
    nl = x.shape[1]
    cp = np.empty_like(x)
    dp = np.empty_like(x)
    inv = T(1) / dd[0]
    cp[:, 0] = inv * u[0]
    dp[:, 0] = inv * r[:, 0]
    for k in range(1, nl):
        inv = T(1) / (dd[k] - a[k] * cp[:, k - 1])
        cp[:, k] = inv * u[k]
        dp[:, k] = inv * (r[:, k] - a[k] * dp[:, k - 1])
    for k in range(nl - 2, -1, -1):
        dp[:, k] = dp[:, k] - cp[:, k] * dp[:, k + 1]
    return x + T(omega) * dp


def restrict(xf, bf, lvf, children):
    T = xf.dtype.type
    r = bf - apply_A(xf, lvf.bands, lvf.cell_next, lvf.sigma)
    acc = np.zeros(r.shape[:2] + (children.shape[0],), dtype=xf.dtype)
    for q in range(children.shape[1]):
        acc = acc + r[..., children[:, q]] * (T(1) / T(4))
    return acc


def vcycle(H, xs, bs, l=0, omega=0.8, nu1=2, nu2=2, ncoarse=8):
    """xs, bs: per-level lists of (nrhs, nl, ncell) arrays; updates xs[l] in place."""
    lv = H.levels[l]
    if l == len(H.levels) - 1:
        for _ in range(ncoarse):
            xs[l][...] = smooth(xs[l], bs[l], lv.bands, lv.cell_next, lv.sigma, omega)
        return
    for _ in range(nu1):
        xs[l][...] = smooth(xs[l], bs[l], lv.bands, lv.cell_next, lv.sigma, omega)
    ch = H.levels[l + 1].children
    bs[l + 1] = restrict(xs[l], bs[l], lv, ch)
    xs[l + 1] = np.zeros_like(bs[l + 1])
    vcycle(H, xs, bs, l + 1, omega, nu1, nu2, ncoarse)
    parent = np.empty(lv.ncell, dtype=np.int64)
    parent[ch.ravel()] = np.repeat(np.arange(ch.shape[0]), ch.shape[1])
    xs[l][...] = xs[l] + xs[l + 1][..., parent]
    for _ in range(nu2):
        xs[l][...] = smooth(xs[l], bs[l], lv.bands, lv.cell_next, lv.sigma, omega)


def to_layer_major(v, ncell, nl):
    """(nrhs, ncell*nl) LFRic W3 layout -> (nrhs, nl, ncell)."""
    return np.ascontiguousarray(v.reshape(v.shape[0], ncell, nl).transpose(0, 2, 1))


def from_layer_major(v):
    return np.ascontiguousarray(v.transpose(0, 2, 1).reshape(v.shape[0], -1))


def solve(H, x, b, ncycles, dtype=np.float64, **kw):
    """Run ncycles V-cycles; x, b in LFRic layout (nrhs, undf).  Returns (x, res)."""
    nl = H.nlayers
    lv = H.levels[0]
    xs = [to_layer_major(x.astype(dtype), lv.ncell, nl)] + [None] * (len(H.levels) - 1)
    bs = [to_layer_major(b.astype(dtype), lv.ncell, nl)] + [None] * (len(H.levels) - 1)
    res = [np.linalg.norm((bs[0] - apply_A(xs[0], lv.bands, lv.cell_next, lv.sigma))
                          .reshape(x.shape[0], -1).astype(np.float64), axis=1)]
    for _ in range(ncycles):
        vcycle(H, xs, bs, 0, **kw)
        res.append(np.linalg.norm((bs[0] - apply_A(xs[0], lv.bands, lv.cell_next, lv.sigma))
                                  .reshape(x.shape[0], -1).astype(np.float64), axis=1))
    return from_layer_major(xs[0]), np.array(res)
