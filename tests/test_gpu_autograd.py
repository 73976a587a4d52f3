"""Differentiation (the C96L70 'differentiation' configuration at small size):
L = <w, u>, u = CG_20(M_W2, D^T p).  Checks: the dot-product test of every
apply/adjoint pair to 1e-12 (FP64), central finite differences (eps = 1e-6)
on 32 random directions for p and for the operator coefficients, and
agreement with the CPU gradient built from the reference Fortran kernels
(oracle/adjoint.py; transpose_matrix_vector forward, dg_matrix_vector adjoint)."""
import numpy as np
import pytest

from lfric_core_amd.configs import precision_config
from lfric_core_amd.rng import Stream

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ITERS = 20


@pytest.fixture(scope="module")
def setup(ctx):
    from lfric_core_amd import api
    cfg = precision_config(4, 6)
    w2, w3 = api.Space(ctx, cfg.w2), api.Space(ctx, cfg.w3)
    st = Stream(2)
    p = st.uniform(cfg.w3.undf)
    w = st.normal(cfg.w2.undf)
    return cfg, w2, w3, p, w


def _ops(ctx, cfg, w2, w3, dt):
    from lfric_core_amd import api
    dev = ctx.device
    M = torch.from_numpy(cfg.mass.astype(dt)).to(dev)
    D = torch.from_numpy(cfg.div.astype(dt)).to(dev)
    # div maps W2 -> W3: local_stencil(ncell_3d, 1, 6), to = W3, from = W2
    return (api.LMAOperator(M, w2, w2, cfg.colouring), api.LMAOperator(D, w3, w2, cfg.colouring))


def _loss(M, D, p, w, mop, dop, iters=ITERS):
    from lfric_core_amd import autograd as ag
    b = ag.lma_apply_transpose(D, p, dop)           # D^T p: W3 -> W2
    u = ag.cg_solve(M, b, mop, iters)
    return torch.dot(w, u)


def test_dot_product_tests(ctx, setup):
    from lfric_core_amd import autograd as ag
    cfg, w2, w3, p, w = setup
    mop, dop = _ops(ctx, cfg, w2, w3, np.float64)
    st = Stream(40)
    dev = ctx.device
    for op, nin, nout in ((mop, cfg.w2.undf, cfg.w2.undf), (dop, cfg.w2.undf, cfg.w3.undf)):
        # <A x, y> on the to-space, <x, A^T y> on the from-space
        x = torch.from_numpy(st.uniform(nin)).to(dev)
        y = torch.from_numpy(st.uniform(nout)).to(dev)
        Ax = ag._apply(op.matrix, x, op)
        ATy = ag._apply_t(op.matrix, y, op)
        a, b = float(torch.dot(Ax, y)), float(torch.dot(x, ATy))
        assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def test_gradient_matches_reference_kernels(ctx, setup):
    import oracle
    from oracle.adjoint import loss_and_grad
    cfg, w2, w3, p, w = setup
    mop, dop = _ops(ctx, cfg, w2, w3, np.float64)
    dev = ctx.device
    L_ref, pbar_ref, Mbar_ref = loss_and_grad(cfg, p, w, ITERS)
    M = mop.matrix.clone().requires_grad_(True)
    pt = torch.from_numpy(p).to(dev).requires_grad_(True)
    L = _loss(M, dop.matrix, pt, torch.from_numpy(w).to(dev), mop, dop)
    L.backward()
    assert abs(float(L) - L_ref) <= 1e-12 * abs(L_ref)
    g = pt.grad.cpu().numpy()
    assert np.linalg.norm(g - pbar_ref) <= 1e-10 * np.linalg.norm(pbar_ref)
    gm = M.grad.cpu().numpy()
    assert np.linalg.norm(gm - Mbar_ref) <= 1e-10 * np.linalg.norm(Mbar_ref)


def test_central_finite_differences_32_directions(ctx, setup):
    cfg, w2, w3, p, w = setup
    mop, dop = _ops(ctx, cfg, w2, w3, np.float64)
    dev = ctx.device
    wt = torch.from_numpy(w).to(dev)
    M = mop.matrix.clone().requires_grad_(True)
    pt = torch.from_numpy(p).to(dev).requires_grad_(True)
    L = _loss(M, dop.matrix, pt, wt, mop, dop)
    L.backward()
    gp, gM = pt.grad.clone(), M.grad.clone()
    st = Stream(41)
    eps = 1e-6
    worst = 0.0
    with torch.no_grad():
        for n in range(32):
            d = torch.from_numpy(st.uniform(cfg.w3.undf)).to(dev)
            dM = torch.from_numpy(st.uniform(M.numel())).to(dev).reshape(M.shape)
            fd_p = (_loss(M, dop.matrix, pt + eps * d, wt, mop, dop)
                    - _loss(M, dop.matrix, pt - eps * d, wt, mop, dop)) / (2 * eps)
            fd_M = (_loss(M + eps * dM, dop.matrix, pt, wt, mop, dop)
                    - _loss(M - eps * dM, dop.matrix, pt, wt, mop, dop)) / (2 * eps)
            an_p = torch.dot(gp, d)
            an_M = torch.sum(gM * dM)
            for fd, an in ((fd_p, an_p), (fd_M, an_M)):
                rel = abs(float(fd - an)) / max(abs(float(an)), 1e-30)
                worst = max(worst, rel)
    print(f"worst FD relative error over 32 directions: {worst:.2e}", flush=True)
    assert worst <= 1e-6


def test_fp32_gradient_close_to_fp64(ctx, setup):
    cfg, w2, w3, p, w = setup
    grads = {}
    for dt in (np.float32, np.float64):
        mop, dop = _ops(ctx, cfg, w2, w3, dt)
        tdt = mop.matrix.dtype
        pt = torch.from_numpy(p).to(ctx.device, tdt).requires_grad_(True)
        L = _loss(mop.matrix, dop.matrix, pt, torch.from_numpy(w).to(ctx.device, tdt), mop, dop)
        L.backward()
        grads[dt] = pt.grad.double().cpu().numpy()
    err = np.linalg.norm(grads[np.float32] - grads[np.float64]) / np.linalg.norm(grads[np.float64])
    print(f"FP32 vs FP64 gradient relative error: {err:.2e}", flush=True)
    assert err <= 1e-3
