"""The numpy restatements of cg_solve / gmres_solve pinned by closed forms
(no reference test covers sci_iterative_solver_mod.f90)."""
import numpy as np
import pytest

from oracle.solvers import cg_solve, gmres_solve
from lfric_core_amd.rng import Stream


def _spd(n, seed):
    st = Stream(seed)
    B = st.normal(n * n).reshape(n, n)
    return B @ B.T + n * np.eye(n)


def test_cg_exact_in_n_iterations():
    n = 12
    A = _spd(n, 1)
    b = Stream(2).uniform(n)
    x = np.zeros(n)
    it, conv, r0, rn = cg_solve(lambda v: A @ v, x, b, r_tol=1e-12, max_iter=n)
    assert conv and it <= n
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def test_cg_directions_are_A_orthogonal():
    n = 20
    A = _spd(n, 3)
    b = Stream(4).uniform(n)
    rec = []
    cg_solve(lambda v: A @ v, np.zeros(n), b, r_tol=1e-14, max_iter=8, record=rec)
    P = np.array(rec)
    G = P @ A @ P.T
    off = G - np.diag(np.diag(G))
    assert np.abs(off).max() <= 1e-9 * np.abs(np.diag(G)).max()


def test_cg_zero_rhs_returns_immediately():
    A = _spd(5, 5)
    x = np.zeros(5)
    assert cg_solve(lambda v: A @ v, x, np.zeros(5))[0] == 0


def test_cg_unmonitored_runs_max_iter_and_diag_prec():
    n = 30
    A = _spd(n, 6)
    b = Stream(7).uniform(n)
    x = np.zeros(n)
    it, conv, _, _ = cg_solve(lambda v: A @ v, x, b, monitor=False, max_iter=7)
    assert it == 7 and not conv
    x = np.zeros(n)
    d = np.diag(A).copy()
    it, conv, _, rn = cg_solve(lambda v: A @ v, x, b, prec=lambda v: v / d, r_tol=1e-10, max_iter=200)
    assert conv
    np.testing.assert_allclose(A @ x, b, atol=1e-8)


def test_cg_float32_vectors():
    n = 40
    A = _spd(n, 8).astype(np.float32)
    b = Stream(9).uniform(n).astype(np.float32)
    x = np.zeros(n, np.float32)
    it, conv, _, _ = cg_solve(lambda v: A @ v, x, b, r_tol=1e-5, max_iter=200)
    assert conv and x.dtype == np.float32
    np.testing.assert_allclose(x, np.linalg.solve(A.astype(np.float64), b), rtol=1e-3, atol=1e-5)


def test_gmres_exact_with_full_krylov_space():
    n = 10
    st = Stream(10)
    A = st.normal(n * n).reshape(n, n) + n * np.eye(n)      # nonsymmetric
    b = st.uniform(n)
    x = np.zeros(n)
    it, conv, err = gmres_solve(lambda v: A @ v, x, b, gcrk=n, r_tol=1e-10, max_iter=1)
    assert conv and it == 1
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)


def test_gmres_restarts_monotone_residual():
    n = 60
    st = Stream(11)
    A = st.normal(n * n).reshape(n, n) * 0.3 + 4 * np.eye(n)
    b = st.uniform(n)
    rec = []
    it, conv, err = gmres_solve(lambda v: A @ v, np.zeros(n), b, gcrk=5, r_tol=1e-10,
                                max_iter=60, record=rec)
    assert conv
    assert all(b_ <= a_ * (1 + 1e-12) for a_, b_ in zip(rec, rec[1:]))


def test_oracle_cg_adjoint_matches_finite_differences():
    """The CPU reverse-mode gradient of L = <w, CG_K(M, D^T p)> (reference
    kernels when built) against central differences, for p and for M."""
    from oracle.adjoint import loss_and_grad
    from lfric_core_amd.configs import precision_config
    cfg = precision_config(2, 3)
    st = Stream(2)
    p = st.uniform(cfg.w3.undf)
    w = st.normal(cfg.w2.undf)
    L, gp, gM = loss_and_grad(cfg, p, w, 10)
    eps = 1e-6
    for _ in range(4):
        d = st.uniform(cfg.w3.undf)
        fd = (loss_and_grad(cfg, p + eps * d, w, 10)[0] - loss_and_grad(cfg, p - eps * d, w, 10)[0]) / (2 * eps)
        assert abs(fd - gp @ d) <= 1e-7 * abs(gp @ d)
        dM = st.uniform(cfg.mass.size).reshape(cfg.mass.shape)
        fd = (loss_and_grad(cfg, p, w, 10, M=cfg.mass + eps * dM)[0]
              - loss_and_grad(cfg, p, w, 10, M=cfg.mass - eps * dM)[0]) / (2 * eps)
        assert abs(fd - (gM * dM).sum()) <= 1e-7 * abs((gM * dM).sum())


# --- the other solver types of sci_iterative_solver_mod.f90 -------------------

from oracle.solvers import (_dot, bicgstab_solve, chebyshev_solve, fgmres_solve, gcr_solve,  # noqa: E402
                            jacobi_solve, precondition_only_solve)


def _nonsym(n, seed, scale=0.3, shift=4.0):
    st = Stream(seed)
    return st.normal(n * n).reshape(n, n) * scale + shift * np.eye(n)


def test_field_dot_is_the_sum_of_field_dots_in_order():
    st = Stream(20)
    x, y = st.uniform(17), st.uniform(17)
    a = float(np.dot(x[:5], y[:5]))
    b = float(np.dot(x[5:12], y[5:12]))
    c = float(np.dot(x[12:], y[12:]))
    assert _dot(x, y, [5, 7, 5]) == (0.0 + a + b) + c


def test_bicgstab_converges_to_the_dense_solution():
    n = 40
    A = _nonsym(n, 21)
    b = Stream(22).uniform(n)
    x = np.zeros(n)
    it, conv, sc, err = bicgstab_solve(lambda v: A @ v, x, b, r_tol=1e-12, max_iter=200)
    assert conv and err < 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
    # diagonal preconditioning reaches the same solution
    d = np.diag(A).copy()
    x2 = np.zeros(n)
    it2, conv2, _, _ = bicgstab_solve(lambda v: A @ v, x2, b, prec=lambda v: v / d, r_tol=1e-12, max_iter=200)
    assert conv2
    np.testing.assert_allclose(x2, x, rtol=1e-9, atol=1e-12)


def test_bicgstab_first_step_closed_form():
    """One iteration from x = 0 with P = I: p = r0, alpha = <r0,r0>/<r0,A r0>,
    s = r0 - alpha A r0, omega = <As, s>/<As, As>, x = omega s + alpha r0."""
    n = 6
    A = _nonsym(n, 23, 0.5, 3.0)
    b = Stream(24).uniform(n)
    x = np.zeros(n)
    bicgstab_solve(lambda v: A @ v, x, b, max_iter=1, monitor=False)
    r0 = b
    alpha = (r0 @ r0) / (r0 @ (A @ r0))
    s = r0 - alpha * (A @ r0)
    t = A @ s
    omega = (t @ s) / (t @ t)
    np.testing.assert_allclose(x, omega * s + alpha * r0, rtol=1e-13)


def test_fgmres_exact_in_one_restart_and_adds_the_correction():
    n = 10
    A = _nonsym(n, 25, 1.0, float(n))
    b = Stream(26).uniform(n)
    x0 = Stream(27).uniform(n)
    x = x0.copy()
    it, conv, err = fgmres_solve(lambda v: A @ v, x, b, gcrk=n, r_tol=1e-10, max_iter=1)
    assert conv and it == 1
    # the reference starts the correction from a zero guess: x = x0 + A^{-1} b
    np.testing.assert_allclose(x - x0, np.linalg.solve(A, b), rtol=1e-9)


def test_fgmres_right_preconditioned_restarts():
    n = 50
    A = _nonsym(n, 28)
    b = Stream(29).uniform(n)
    d = np.diag(A).copy()
    x = np.zeros(n)
    it, conv, err = fgmres_solve(lambda v: A @ v, x, b, prec=lambda v: v / d, gcrk=4, r_tol=1e-11, max_iter=100)
    assert conv and it > 1
    np.testing.assert_allclose(A @ x, b, atol=1e-9)


def test_gcr_exact_in_n_steps_orthonormal_directions():
    n = 12
    A = _nonsym(n, 30, 1.0, float(n))
    b = Stream(31).uniform(n)
    x = np.zeros(n)
    rec = []
    steps, conv, err = gcr_solve(lambda v: A @ v, x, b, gcrk=n, r_tol=1e-11, max_iter=1, record=rec)
    assert conv and steps <= n
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)
    V = np.array(rec)
    np.testing.assert_allclose(V @ V.T, np.eye(len(rec)), atol=1e-10)


def test_gcr_initial_guess_is_preconditioned_rhs():
    """With max_iter = 0 GCR returns x0 + P^{-1} b (its starting correction)."""
    n = 8
    A = _nonsym(n, 32)
    b = Stream(33).uniform(n)
    d = np.diag(A).copy()
    x0 = Stream(34).uniform(n)
    x = x0.copy()
    gcr_solve(lambda v: A @ v, x, b, prec=lambda v: v / d, gcrk=3, max_iter=0)
    np.testing.assert_allclose(x, x0 + b / d, rtol=1e-15)


def test_gcr_residuals_monotone_with_restarts():
    n = 40
    A = _nonsym(n, 35)
    b = Stream(36).uniform(n)
    errs = []
    for k in range(1, 8):
        _, _, err = gcr_solve(lambda v: A @ v, np.zeros(n), b, gcrk=3, r_tol=0.0, max_iter=k)
        errs.append(err)
    assert all(e2 <= e1 * (1 + 1e-12) for e1, e2 in zip(errs, errs[1:]))


def test_jacobi_error_propagation_closed_form():
    """e_k = (I - rho D^{-1} A)^k e_0 for the damped Jacobi iteration."""
    n = 15
    A = _nonsym(n, 37, 0.2, 3.0)
    b = Stream(38).uniform(n)
    d = np.diag(A).copy()
    xs = np.linalg.solve(A, b)
    x0 = Stream(39).uniform(n)
    x = x0.copy()
    rho = 0.8
    k = 6
    jacobi_solve(lambda v: A @ v, x, b, prec=lambda v: v / d, max_iter=k, monitor=False, rho_relax=rho)
    G = np.eye(n) - rho * (A / d[:, None])
    np.testing.assert_allclose(x - xs, np.linalg.matrix_power(G, k) @ (x0 - xs), rtol=1e-10, atol=1e-13)


def test_jacobi_converges_monitored():
    n = 30
    A = _nonsym(n, 40, 0.1, 3.0)
    b = Stream(41).uniform(n)
    d = np.diag(A).copy()
    x = np.zeros(n)
    it, conv, r0, rn = jacobi_solve(lambda v: A @ v, x, b, prec=lambda v: v / d, r_tol=1e-10, max_iter=500)
    assert conv and rn / r0 <= 1e-10
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8)


def test_chebyshev_error_polynomial_closed_form():
    """On diagonal A = diag(lam) with P = I and x0 = 0 the reference's
    recurrence (w_0 = 1, w_k = 1/(1 - w_{k-1}/(4 a2^2))) gives
    e_k = p_k(lam) e_0 with, for xi = (lmax + lmin - 2 lam)/(lmax - lmin),
    a1 = 2/(lmax - lmin), a2 = (lmax + lmin)/(lmax - lmin) and U_k the
    Chebyshev polynomials of the second kind,
        p_k = (U_{k+1}(xi) + 2 a1 lam U_k(xi)) / U_{k+1}(a2);
    U is evaluated from its trigonometric / hyperbolic forms."""
    lmin, lmax = 0.5, 4.0
    lam = np.linspace(lmin, lmax, 11)[1:-1]
    b = Stream(42).uniform(lam.size)
    xs = b / lam
    a1 = 2.0 / (lmax - lmin)
    a2 = (lmax + lmin) / (lmax - lmin)
    xi = (lmax + lmin - 2 * lam) / (lmax - lmin)
    th = np.arccos(xi)
    ph = np.arccosh(a2)
    U = lambda k: np.sin((k + 1) * th) / np.sin(th)            # noqa: E731
    Ua = lambda k: np.sinh((k + 1) * ph) / np.sinh(ph)          # noqa: E731
    for k in (1, 2, 5, 9):
        x = np.zeros(lam.size)
        chebyshev_solve(lambda v: lam * v, x, b, max_iter=k, lmin=lmin, lmax=lmax, monitor=False)
        pk = (U(k + 1) + 2 * a1 * lam * U(k)) / Ua(k + 1)
        np.testing.assert_allclose(x - xs, pk * (0 - xs), rtol=1e-9, atol=1e-13)
    # and it converges on the interval
    x = np.zeros(lam.size)
    chebyshev_solve(lambda v: lam * v, x, b, max_iter=40, lmin=lmin, lmax=lmax, monitor=False)
    np.testing.assert_allclose(x, xs, rtol=1e-8)


def test_precondition_only():
    n = 9
    A = _nonsym(n, 43)
    b = Stream(44).uniform(n)
    d = np.diag(A).copy()
    x = np.full(n, 7.0)
    bn, rn = precondition_only_solve(lambda v: A @ v, x, b, prec=lambda v: v / d)
    np.testing.assert_array_equal(x, b / d)
    assert abs(bn - np.linalg.norm(b)) <= 1e-14 * bn
    assert abs(rn - np.linalg.norm(b - A @ x)) <= 1e-13 * bn
