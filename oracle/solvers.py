"""Iterative solvers restated from the reference in numpy (TEST INFRASTRUCTURE ONLY).

Each function follows one solver type of
components/science/source/solver/sci_iterative_solver_mod.f90 statement by
statement, with FP64 scalars: ``cg_solve`` (:495-592), ``bicgstab_solve``
(:645-779), ``gmres_solve`` (:830-1003), ``fgmres_solve`` (:1055-1237),
``gcr_solve`` (:1288-1408), ``precondition_only_solve`` (:1664-1694),
``jacobi_solve`` (:1745-1815) and ``chebyshev_solve`` (:1868-1940).  Vector
storage may be float32 (r_solver), in which case scalars are cast to float32
before every update and inner products accumulate in float64, as
r_solver_field_vector_type does.  ``fields`` (a list of field lengths) makes
the vectors n-field vectors: inner products and norms are then the sum, in
field order, of the per-field inner products, as field_vector_type%dot and
%norm form them (components/science/source/algorithm/solver/
sci_field_vector_mod.x90:59-107).

Parity: no reference test exercises these solvers, so they are pinned by
closed forms (tests/test_solvers_oracle.py): agreement with a dense direct
solve; exact convergence of CG on an n x n SPD system in n iterations, of
GMRES and FGMRES with gcrk >= n in one restart and of GCR in n steps;
A-orthogonality of the CG search directions and of the GCR directions;
monotone GMRES and GCR residuals; BiCGstab on a symmetric system against
exact arithmetic of a 2 x 2 case; the Jacobi fixed point and rate; the
Chebyshev error polynomial against its closed form on a diagonal system.
"""

from __future__ import annotations

import numpy as np

EPS = np.finfo(np.float64).eps      # constants_mod.F90:167


def _dot(x, y, fields=None):
    if fields is None:
        return float(np.dot(x.astype(np.float64), y.astype(np.float64)))
    tot, o = 0.0, 0
    for n in fields:
        tot = tot + float(np.dot(x[o:o + n].astype(np.float64), y[o:o + n].astype(np.float64)))
        o += n
    return tot


def _norm(x, fields=None):
    return float(np.sqrt(_dot(x, x, fields)))


def cg_solve(apply, x, b, prec=None, r_tol=1e-8, a_tol=0.0, max_iter=100, monitor=True,
             record=None, fields=None):
    """Solve A x = b in place.  apply(v) -> A v; prec(v) -> P^{-1} v (None: copy).
    Returns (iterations, converged, r_nrm0, r_nrm)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    r = apply(x)
    r = T(-1.0) * r
    r = r + T(1.0) * b
    r_nrm_0 = np.sqrt(_dot(r, r, fields))
    if r_nrm_0 < EPS:
        return 0, True, r_nrm_0, r_nrm_0
    z = prec(r)
    rz = _dot(r, z, fields)
    r_nrm = r_nrm_0
    p = z.copy()
    converged = False
    it = 0
    for it in range(1, max_iter + 1):
        z = apply(p)
        if record is not None:
            record.append(p.copy())
        alpha = rz / _dot(p, z, fields)
        x += T(alpha) * p
        r += T(-alpha) * z
        if monitor:
            r_nrm = np.sqrt(_dot(r, r, fields))
            if r_nrm / r_nrm_0 <= r_tol or r_nrm <= a_tol:
                converged = True
                break
        z = prec(r)
        rz_new = _dot(r, z, fields)
        beta = rz_new / rz
        p = T(beta) * p + z
        rz = rz_new
    if not monitor:
        r_nrm = np.sqrt(_dot(r, r, fields))
    return it, converged, r_nrm_0, r_nrm


def gmres_solve(apply, x, b, prec=None, gcrk=10, r_tol=1e-8, a_tol=0.0, max_iter=10,
                monitor=True, record=None, fields=None):
    """Left-preconditioned restarted GMRES(gcrk) in place.  Returns
    (restarts, converged, err) with err = ||b - A x|| / max(||r0||, a_tol)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    Ax = apply(x)
    res = b.copy()
    res = res + T(-1.0) * Ax
    sc_err = np.sqrt(_dot(res, res, fields))
    if sc_err < EPS:
        return 0, True, 0.0
    if monitor:
        sc_err = max(sc_err, a_tol)
    err = np.inf
    it = 0
    for it in range(1, max_iter + 1):
        s = prec(res)
        beta = np.sqrt(_dot(s, s, fields))
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
                h[ivj, iv] = _dot(v[ivj], w, fields)
                w = w + T(-h[ivj, iv]) * v[ivj]
            h[iv + 1, iv] = np.sqrt(_dot(w, w, fields))
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
            err = np.sqrt(_dot(res, res, fields)) / sc_err
            if record is not None:
                record.append(err)
            if err < r_tol:
                return it, True, err
    return it, False, err


def _givens_back(h, g, gcrk):
    """The Givens rotations and back substitution of gmres_solve / fgmres_solve."""
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
    return u


def bicgstab_solve(apply, x, b, prec=None, r_tol=1e-8, a_tol=0.0, max_iter=100, monitor=True,
                   fields=None):
    """BiCGstab in place.  Returns (iterations, converged, sc_err, err)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    v = apply(x)
    r = b + T(-1.0) * v
    r0 = r.copy()
    sc_err = _norm(r, fields)
    if sc_err < EPS:
        return 0, True, sc_err, 0.0
    sc_err = max(sc_err, a_tol)
    alpha = omega = rho_old = 1.0
    v = np.zeros_like(x)
    p = np.zeros_like(x)
    err = np.inf
    it = 0
    for it in range(1, max_iter + 1):
        rho = _dot(r, r0, fields)
        beta = (rho / rho_old) * (alpha / omega)
        t = r + T(-(beta * omega)) * v
        y = prec(t)
        p = T(beta) * p + y
        v = apply(p)
        alpha = rho / _dot(r0, v, fields)
        s = r + T(-alpha) * v
        z = prec(s)
        t = apply(z)
        tt = _dot(t, t, fields)
        ts = _dot(t, s, fields)
        omega = ts / tt
        x += T(omega) * z
        x += T(alpha) * p
        r = s + T(-omega) * t
        rho_old = rho
        if monitor:
            err = _norm(r, fields) / sc_err
            if err < r_tol:
                return it, True, sc_err, err
    return it, False, sc_err, err


def fgmres_solve(apply, x, b, prec=None, gcrk=10, r_tol=1e-8, a_tol=0.0, max_iter=10, monitor=True,
                 fields=None):
    """Flexible GMRES: the correction dx starts from zero (residual b) and is
    added to x at the end.  Returns (restarts, converged, err)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    dx = np.zeros_like(x)
    res = b.copy()
    sc_err = _norm(res, fields)
    if sc_err < EPS:
        return 0, True, 0.0
    if monitor:
        sc_err = max(sc_err, a_tol)
    beta = _norm(res, fields)
    v = [None] * gcrk
    Pv = [None] * gcrk
    v[0] = T(1.0 / beta) * res
    g = np.zeros(gcrk + 1)
    g[0] = beta
    err = np.inf
    it = 0
    for it in range(1, max_iter + 1):
        h = np.zeros((gcrk + 1, gcrk))
        for iv in range(gcrk):
            Pv[iv] = prec(v[iv])
            w = apply(Pv[iv])
            for ivj in range(iv + 1):
                h[ivj, iv] = _dot(v[ivj], w, fields)
                w = w + T(-h[ivj, iv]) * v[ivj]
            h[iv + 1, iv] = _norm(w, fields)
            if iv < gcrk - 1:
                v[iv + 1] = T(1.0 / h[iv + 1, iv]) * w
        u = _givens_back(h, g, gcrk)
        for iv in range(gcrk):
            dx += T(u[iv]) * Pv[iv]
        Ax = apply(dx)
        res = T(-1.0) * Ax + b
        beta = _norm(res, fields)
        if monitor:
            err = beta / sc_err
            if err < r_tol:
                x += T(1.0) * dx
                return it, True, err
        v[0] = T(1.0 / beta) * res
        g = np.zeros(gcrk + 1)
        g[0] = beta
    x += T(1.0) * dx
    return it, False, err


def gcr_solve(apply, x, b, prec=None, gcrk=10, r_tol=1e-8, a_tol=0.0, max_iter=10, monitor=True,
              fields=None, record=None):
    """Restarted GCR: the correction dx starts from P^{-1} b and is added to x
    at the end.  Returns (inner steps, converged, err)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    dx = prec(b)
    Ax = apply(dx)
    res = b + T(-1.0) * Ax
    sc_err = _norm(b, fields)
    if sc_err < EPS:
        return 0, True, 0.0
    if monitor:
        sc_err = max(sc_err, a_tol)
    v = [None] * gcrk
    Pv = [None] * gcrk
    err = np.inf
    steps = 0
    for _it in range(1, max_iter + 1):
        for iv in range(gcrk):
            Pv[iv] = prec(res)
            v[iv] = apply(Pv[iv])
            for ivj in range(iv):
                alpha = _dot(v[iv], v[ivj], fields)
                v[iv] = v[iv] + T(-alpha) * v[ivj]
                Pv[iv] = Pv[iv] + T(-alpha) * Pv[ivj]
            alpha = _norm(v[iv], fields)
            beta = 1.0 / alpha
            v[iv] = T(beta) * v[iv]
            Pv[iv] = T(beta) * Pv[iv]
            alpha = _dot(res, v[iv], fields)
            dx += T(alpha) * Pv[iv]
            res = res + T(-alpha) * v[iv]
            steps += 1
            if record is not None:
                record.append(v[iv].copy())
            if monitor:
                err = _norm(res, fields) / sc_err
                if err < r_tol:
                    x += T(1.0) * dx
                    return steps, True, err
    x += T(1.0) * dx
    return steps, False, err


def precondition_only_solve(apply, x, b, prec=None, fields=None):
    """x = P^{-1} b; returns (||b||, ||b - A x||)."""
    prec = prec or (lambda v: v.copy())
    x[:] = prec(b)
    res = b + x.dtype.type(-1.0) * apply(x)
    return _norm(b, fields), _norm(res, fields)


def jacobi_solve(apply, x, b, prec=None, r_tol=1e-8, a_tol=0.0, max_iter=100, monitor=True, rho_relax=1.0,
                 fields=None):
    """x -= rho_relax P^{-1}(A x - b), max_iter times or to convergence.
    Returns (iterations, converged, r_nrm0, r_nrm)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    const = -rho_relax
    r_nrm0 = r_nrm = 0.0
    it = 0
    for it in range(1, max_iter + 1):
        r = apply(x)
        r = r + T(-1.0) * b
        if monitor:
            r_nrm = _norm(r, fields)
            if it == 1:
                r_nrm0 = r_nrm
            if r_nrm / r_nrm0 <= r_tol or r_nrm <= a_tol:
                return it, True, r_nrm0, r_nrm
        z = prec(r)
        x += T(const) * z
    return it, False, r_nrm0, r_nrm


def chebyshev_solve(apply, x, b, prec=None, max_iter=10, lmin=0.1, lmax=1.0, monitor=True, fields=None):
    """max_iter Chebyshev steps from x = 0.  Returns (init_norm, final_norm)."""
    T = x.dtype.type
    prec = prec or (lambda v: v.copy())
    x[:] = 0
    xp = np.zeros_like(x)
    xo = np.zeros_like(x)
    init_norm = max(1.0, _norm(b, fields)) if monitor else 0.0
    a1 = 2.0 / (lmax - lmin)
    a2 = (lmax + lmin) / (lmax - lmin)
    a_over_b = a1 / a2
    w = 1.0
    for it in range(1, max_iter + 1):
        z = apply(xo)
        r = T(1.0) * b + T(-1.0) * z
        z = prec(r)
        w = 1.0 / (1.0 - w / (4.0 * a2 ** 2))
        wab = w * a_over_b
        x[:] = T(wab) * z + T(w) * xo
        x += T(1.0 - w) * xp
        if it < max_iter:
            xp = xo.copy()
            xo = x.copy()
    final = 0.0
    if monitor:
        z = apply(x)
        z = z + T(-1.0) * b
        final = _norm(z, fields)
    return init_norm, final
