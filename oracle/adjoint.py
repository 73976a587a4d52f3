"""Reverse-mode gradient of L = <w, CG_K(M, D^T p)> on the CPU (TEST
INFRASTRUCTURE ONLY).

The forward operators are the oracle's kernels - or, when oracle/_ref is
built, the reference's own Fortran kernels: b = D^T p through
transpose_matrix_vector (components/science/source/kernel/algebra/
transpose_matrix_vector_kernel_mod.F90) and the CG applies through
matrix_vector.  The adjoint of D^T is D applied with dg_matrix_vector.  The
CG reverse sweep stores every iterate (no checkpointing), independently of
the library's checkpointed sweep, and follows the statement order of
oracle.solvers.cg_solve.
"""

from __future__ import annotations

import numpy as np

import oracle


def _dot(a, b):
    return float(np.dot(a.astype(np.float64), b.astype(np.float64)))


def cg_forward_record(apply, b, iters):
    T = b.dtype.type
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rz = _dot(r, r)
    hist = []
    for _ in range(iters):
        q = apply(p)
        pq = _dot(p, q)
        alpha = rz / pq
        x = x + T(alpha) * p
        r_next = r + T(-alpha) * q
        rz_next = _dot(r_next, r_next)
        beta = rz_next / rz
        hist.append((p, q, r_next, pq, alpha, rz, rz_next, beta))
        p = T(beta) * p + r_next
        r, rz = r_next, rz_next
    return x, hist


def cg_reverse(apply_t, b, hist, xbar, want_outer=None):
    """Returns (bbar, sum_i qbar_i (x) p_i terms via want_outer callback)."""
    T = b.dtype.type
    rbar = np.zeros_like(b)
    pbar = np.zeros_like(b)
    rzbar = 0.0
    for (p_i, q_i, r_next, pq, alpha, rz_i, rz_next, beta) in reversed(hist):
        rbar = rbar + T(1.0) * pbar
        betabar = _dot(pbar, p_i)
        pbar = T(beta) * pbar
        rzbar_next = rzbar + betabar / rz_i
        rzbar_i = -betabar * rz_next / (rz_i * rz_i)
        rbar = rbar + T(2.0 * rzbar_next) * r_next
        alphabar = -_dot(rbar, q_i) + _dot(xbar, p_i)
        qbar = T(-alpha) * rbar
        pbar = pbar + T(alpha) * xbar
        rzbar_i += alphabar / pq
        pqbar = -alphabar * rz_i / (pq * pq)
        pbar = pbar + T(pqbar) * q_i
        qbar = qbar + T(pqbar) * p_i
        pbar = pbar + apply_t(qbar)
        if want_outer is not None:
            want_outer(qbar, p_i)
        rzbar = rzbar_i
    rbar = rbar + pbar
    rbar = rbar + T(2.0 * rzbar) * b
    return rbar


def loss_and_grad(cfg, p, w, iters=20, M=None, D=None, use_ref=None, dtype=np.float64, M_dirs=None):
    """L = <w, CG_iters(M, D^T p)> and dL/dp, dL/dM on the CPU.

    With ``M_dirs`` (a list of arrays shaped like M) the third result is the
    list of directional derivatives <dL/dM, dM> instead of dL/dM itself,
    formed as sum_i <qbar_i, dM p_i> with one apply of dM per iteration (the
    matrix gradient is sum_i qbar_i (x) p_i per cell-layer)."""
    use_ref = oracle.ref() is not None if use_ref is None else use_ref
    M = (cfg.mass if M is None else M).astype(dtype)
    D = (cfg.div if D is None else D).astype(dtype)
    p = p.astype(dtype)
    w = w.astype(dtype)
    nl = cfg.nlayers
    dm2, dm3 = cfg.w2.dofmap, cfg.w3.dofmap
    col = cfg.colouring

    def A(v):
        return oracle.invoke_coloured(0, col, nl, np.zeros_like(v), v, M, dm2, dm2, use_ref=use_ref)

    def AT(v):
        return oracle.invoke_coloured(1, col, nl, np.zeros_like(v), v, M, dm2, dm2, use_ref=use_ref)

    b = oracle.invoke_coloured(1, col, nl, np.zeros(cfg.w2.undf, dtype), p, D, dm2, dm3,
                               use_ref=use_ref)
    u, hist = cg_forward_record(A, b, iters)
    L = _dot(w, u)
    if M_dirs is not None:
        proj = [0.0] * len(M_dirs)

        def outer(qbar, p_i):
            for q, dM in enumerate(M_dirs):
                dMp = oracle.invoke_coloured(0, col, nl, np.zeros_like(p_i), p_i, dM.astype(dtype), dm2, dm2,
                                             use_ref=use_ref)
                proj[q] += _dot(qbar, dMp)

        bbar = cg_reverse(AT, b, hist, w, outer)
        pbar = oracle.invoke_dg(0, nl, np.zeros(cfg.w3.undf, dtype), bbar, D, dm3, dm2, use_ref=use_ref)
        return L, pbar, proj
    Mbar = np.zeros_like(M)
    # dof of (cell-layer ik, local dof) on W2
    idx = (dm2[:, None, :] - 1 + np.arange(nl)[None, :, None]).reshape(-1, dm2.shape[1])

    def outer(qbar, p_i):
        # Mbar[j, i, ik] += qbar(map(i)+k) * p_i(map(j)+k)
        Mbar[...] += np.einsum("ki,kj->jik", qbar[idx], p_i[idx])

    bbar = cg_reverse(AT, b, hist, w, outer)
    pbar = oracle.invoke_dg(0, nl, np.zeros(cfg.w3.undf, dtype), bbar, D, dm3, dm2,
                            use_ref=use_ref)
    return L, pbar, Mbar
