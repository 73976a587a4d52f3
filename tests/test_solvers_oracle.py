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
