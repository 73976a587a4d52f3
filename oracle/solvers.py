"""Krylov solvers restated from the reference in numpy (TEST INFRASTRUCTURE ONLY).

``cg_solve`` follows conjugate_gradient_type%cg_solve and ``gmres_solve``
follows gmres_type%gmres_solve
(components/science/source/solver/sci_iterative_solver_mod.f90:495-592 and
:830-1003) statement by statement, with FP64 scalars; vector storage may be
float32 (r_solver) in which case scalars are cast to float32 before every
update and inner products accumulate in float64, as
r_solver_field_vector_type does.

Parity: no reference test exercises these solvers, so they are pinned by
closed forms (tests/test_solvers_oracle.py): agreement with a dense direct
solve, exact convergence of CG on an n x n SPD system in n iterations and of
GMRES with gcrk >= n in one restart, A-orthogonality of the CG search
directions and monotone GMRES residuals.
"""

from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).eps      # constants_mod.F90:167


def _dot(x, y):
    return float(np.dot(x.astype(np.float64), y.astype(np.float64)))


def cg_solve(apply, x, b, prec=None, r_tol=1e-8, a_tol=0.0, max_iter=100, monitor=True,
             record=None):
    """Solve A x = b in place.  apply(v) -> A v; prec(v) -> P^{-1} v (None: copy).
    Returns (iterations, converged, r_nrm0, r_nrm)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    r = apply(x)
    r = T(-1.0) * r
    r = r + T(1.0) * b
    r_nrm_0 = np.sqrt(_dot(r, r))
    if r_nrm_0 < EPS:
        return 0, True, r_nrm_0, r_nrm_0
    z = prec(r)
    rz = _dot(r, z)
    r_nrm = r_nrm_0
    p = z.copy()
    converged = False
    it = 0
    for it in range(1, max_iter + 1):
        z = apply(p)
        if record is not None:
            record.append(p.copy())
        alpha = rz / _dot(p, z)
        x += T(alpha) * p
        r += T(-alpha) * z
        if monitor:
            r_nrm = np.sqrt(_dot(r, r))
            if r_nrm / r_nrm_0 <= r_tol or r_nrm <= a_tol:
                converged = True
                break
        z = prec(r)
        rz_new = _dot(r, z)
        beta = rz_new / rz
        p = T(beta) * p + z
        rz = rz_new
    if not monitor:
        r_nrm = np.sqrt(_dot(r, r))
    return it, converged, r_nrm_0, r_nrm


def gmres_solve(apply, x, b, prec=None, gcrk=10, r_tol=1e-8, a_tol=0.0, max_iter=10,
                monitor=True, record=None):
    """Left-preconditioned restarted GMRES(gcrk) in place.  Returns
    (restarts, converged, err) with err = ||b - A x|| / max(||r0||, a_tol)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    Ax = apply(x)
    res = b.copy()
    res = res + T(-1.0) * Ax
    sc_err = np.sqrt(_dot(res, res))
    if sc_err < EPS:
        return 0, True, 0.0
    if monitor:
        sc_err = max(sc_err, a_tol)
    err = np.inf
    it = 0
    for it in range(1, max_iter + 1):
        s = prec(res)
        beta = np.sqrt(_dot(s, s))
        v = [None] * gcrk
        v[0] = T(1.0 / beta) * s
        h = np.zeros((gcrk + 1, gcrk))
        g = np.zeros(gcrk + 1)
        g[0] = beta
        for iv in range(gcrk):
            w = v[iv].copy()
            s = apply(w)
            w = prec(s)
            for ivj in range(iv + 1):
                h[ivj, iv] = _dot(v[ivj], w)
                w = w + T(-h[ivj, iv]) * v[ivj]
            h[iv + 1, iv] = np.sqrt(_dot(w, w))
            if iv < gcrk - 1:
                v[iv + 1] = T(1.0 / h[iv + 1, iv]) * w
        for iv in range(gcrk):
            nrm = np.sqrt(h[iv, iv] * h[iv, iv] + h[iv + 1, iv] * h[iv + 1, iv])
            si = h[iv + 1, iv] / nrm
            ci = h[iv, iv] / nrm
            p_ = ci * g[iv] + si * g[iv + 1]
            q_ = -si * g[iv] + ci * g[iv + 1]
            g[iv], g[iv + 1] = p_, q_
            for ivj in range(iv, gcrk):
                h1 = ci * h[iv, ivj] + si * h[iv + 1, ivj]
                h2 = -si * h[iv, ivj] + ci * h[iv + 1, ivj]
                h[iv, ivj], h[iv + 1, ivj] = h1, h2
        u = np.zeros(gcrk)
        u[gcrk - 1] = g[gcrk - 1] / h[gcrk - 1, gcrk - 1]
        for iv in range(gcrk - 2, -1, -1):
            u[iv] = g[iv]
            for ivj in range(iv + 1, gcrk):
                u[iv] = u[iv] - h[iv, ivj] * u[ivj]
            u[iv] = u[iv] / h[iv, iv]
        for iv in range(gcrk):
            x += T(u[iv]) * v[iv]
        Ax = apply(x)
        res = T(-1.0) * Ax + b
        if monitor:
            err = np.sqrt(_dot(res, res)) / sc_err
            if record is not None:
                record.append(err)
            if err < r_tol:
                return it, True, err
    return it, False, err
