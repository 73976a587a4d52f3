"""GPU parity of the LMA applies against the CPU oracle (bit-exact).

Every case builds a cubed-sphere problem, runs the HIP path through the C ABI
and the oracle's coloured / cell loops on the same bytes, and requires
identical results: the HIP kernels keep the reference's summation order and
form each term as a separate multiply and add.
"""

import numpy as np
import pytest

import oracle
from lfric_core_amd.function_space import build_dofmap, colour_cells
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream
from lfric_core_amd.configs import normal_lma, spd_lma

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _setup(ctx, nc, nl, k, spaces):
    from lfric_core_amd import api
    m = cubed_sphere(nc)
    col = colour_cells(m)
    fs = {s: build_dofmap(m, s, nl, k) for s in spaces}
    sp = {s: api.Space(ctx, f) for s, f in fs.items()}
    return m, col, fs, sp


def _t(a, ctx):
    return torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)


CASES = [
    # (nc, nl, k, to, from)
    (4, 5, 0, "W2", "W2"),
    (6, 10, 0, "W0", "W0"),
    (6, 10, 0, "W2", "W3"),
    (5, 7, 0, "W2", "W2"),      # odd panel size: the other colour pattern
    (3, 4, 1, "W2", "W2"),
    (3, 4, 1, "W0", "W0"),
    (3, 4, 1, "W0", "W3"),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl,k,to,frm", CASES)
def test_matrix_vector_bitexact(ctx, dtype, nc, nl, k, to, frm):
    from lfric_core_amd import api
    print(f"case matrix_vector nc={nc} nl={nl} k={k} {to}<-{frm} {dtype.__name__}", flush=True)
    m, col, fs, sp = _setup(ctx, nc, nl, k, {to, frm})
    st = Stream(11)
    n3 = m.ncells * nl
    M = normal_lma(st, n3, fs[to].ndf, fs[frm].ndf).astype(dtype)
    x = st.uniform(fs[frm].undf).astype(dtype)
    y0 = st.uniform(fs[to].undf).astype(dtype)
    # oracle: GH_INC from y0, and overwrite (setval_c(0) then matrix_vector)
    ref_inc = oracle.invoke_coloured(0, col, nl, y0.copy(), x, M, fs[to].dofmap, fs[frm].dofmap)
    ref_ovw = oracle.invoke_coloured(0, col, nl, np.zeros_like(y0), x, M, fs[to].dofmap, fs[frm].dofmap)
    op = api.LMAOperator(_t(M, ctx), sp[to], sp[frm], col)
    y = _t(y0, ctx)
    api.matrix_vector(y, _t(x, ctx), op)
    z = torch.full_like(y, float("nan"))
    api.matrix_vector(z, _t(x, ctx), op, overwrite=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref_inc)
    np.testing.assert_array_equal(z.cpu().numpy(), ref_ovw)


TCASES = [
    (4, 5, 0, "W3", "W2"),   # D^T: adjoint of div, output continuous W2
    (6, 10, 0, "W2", "W2"),
    (6, 6, 0, "W0", "W0"),
    (3, 4, 1, "W3", "W2"),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl,k,to,frm", TCASES)
def test_transpose_bitexact(ctx, dtype, nc, nl, k, to, frm):
    """lhs(frm) += M^T x(to) with M = local_stencil(n3, ndf_to, ndf_frm)."""
    from lfric_core_amd import api
    print(f"case transpose nc={nc} nl={nl} k={k} {frm}<-{to}^T {dtype.__name__}", flush=True)
    m, col, fs, sp = _setup(ctx, nc, nl, k, {to, frm})
    st = Stream(12)
    n3 = m.ncells * nl
    M = normal_lma(st, n3, fs[to].ndf, fs[frm].ndf).astype(dtype)
    xt = st.uniform(fs[to].undf).astype(dtype)
    l0 = st.uniform(fs[frm].undf).astype(dtype)
    ref = oracle.invoke_coloured(1, col, nl, l0.copy(), xt, M, fs[frm].dofmap, fs[to].dofmap)
    op = api.LMAOperator(_t(M, ctx), sp[to], sp[frm], col)
    lhs = _t(l0, ctx)
    api.transpose_matrix_vector(lhs, _t(xt, ctx), op)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(lhs.cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl,k,to,frm", [(6, 10, 0, "W3", "W2"), (6, 10, 0, "W3", "W3"),
                                            (4, 7, 0, "W3", "W2"), (4, 7, 0, "W3", "W3"),   # odd nl
                                            (3, 4, 1, "W3", "W2"), (3, 4, 1, "W3", "W3")])
def test_dg_bitexact(ctx, dtype, nc, nl, k, to, frm):
    from lfric_core_amd import api
    print(f"case dg nc={nc} nl={nl} k={k} {to}<-{frm} {dtype.__name__}", flush=True)
    m, col, fs, sp = _setup(ctx, nc, nl, k, {to, frm})
    st = Stream(13)
    n3 = m.ncells * nl
    M = normal_lma(st, n3, fs[to].ndf, fs[frm].ndf).astype(dtype)
    x = st.uniform(fs[frm].undf).astype(dtype)
    y0 = st.uniform(fs[to].undf).astype(dtype)
    ref_w = oracle.invoke_dg(0, nl, y0.copy(), x, M, fs[to].dofmap, fs[frm].dofmap)
    ref_i = oracle.invoke_dg(1, nl, y0.copy(), x, M, fs[to].dofmap, fs[frm].dofmap)
    op = api.LMAOperator(_t(M, ctx), sp[to], sp[frm], col)
    y = _t(y0, ctx)
    api.dg_matrix_vector(y, _t(x, ctx), op)
    z = _t(y0, ctx)
    api.dg_inc_matrix_vector(z, _t(x, ctx), op)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref_w)
    np.testing.assert_array_equal(z.cpu().numpy(), ref_i)


def test_dot_product_identity_fp64(ctx):
    """<A x, y> = <x, A^T y> to 1e-12 for the W2 mass and div operators."""
    from lfric_core_amd import api
    m, col, fs, sp = _setup(ctx, 12, 20, 0, {"W2", "W3"})
    st = Stream(14)
    n3 = m.ncells * 20
    for to, frm in (("W2", "W2"), ("W3", "W2"), ("W2", "W3")):
        M = normal_lma(st, n3, fs[to].ndf, fs[frm].ndf)
        op = api.LMAOperator(_t(M, ctx), sp[to], sp[frm], col)
        x = _t(st.uniform(fs[frm].undf), ctx)
        y = _t(st.uniform(fs[to].undf), ctx)
        Ax = torch.zeros(fs[to].undf, dtype=torch.float64, device=ctx.device)
        if fs[to].is_discontinuous:
            api.dg_matrix_vector(Ax, x, op)
        else:
            api.matrix_vector(Ax, x, op, overwrite=True)
        ATy = torch.zeros(fs[frm].undf, dtype=torch.float64, device=ctx.device)
        api.transpose_matrix_vector(ATy, y, op)
        a = api.dot(ctx, Ax, y)
        b = api.dot(ctx, x, ATy)
        assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)), (to, frm, a, b)


def test_vector_ops_and_dot(ctx):
    from lfric_core_amd import api
    st = Stream(15)
    n = 100_003
    for dt, tdt in ((np.float32, torch.float32), (np.float64, torch.float64)):
        x = st.uniform(n).astype(dt)
        y = st.uniform(n).astype(dt)
        X, Y = _t(x, ctx), _t(y, ctx)
        got = api.dot(ctx, X, Y)
        want = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
        assert abs(got - want) <= 1e-12 * abs(want) + 1e-12
        assert api.dot(ctx, X, Y) == got          # deterministic
        a = 0.37
        Z = Y.clone()
        api.axpy(ctx, a, X, Z)
        np.testing.assert_array_equal(Z.cpu().numpy(), y + dt(a) * x)
        Z = Y.clone()
        api.aypx(ctx, a, X, Z)
        np.testing.assert_array_equal(Z.cpu().numpy(), dt(a) * y + x)
        W = torch.empty_like(X)
        api.axpby(ctx, a, X, -1.5, Y, W)
        np.testing.assert_array_equal(W.cpu().numpy(), dt(a) * x + dt(-1.5) * y)
    x64 = _t(st.uniform(n), ctx)
    x32 = torch.empty(n, dtype=torch.float32, device=ctx.device)
    api.convert(ctx, x32, x64)
    np.testing.assert_array_equal(x32.cpu().numpy(), x64.cpu().numpy().astype(np.float32))


def test_w2_mass_c96l70_fp32_bitexact(ctx):
    """Full-size W2 mass apply (C96, 70 layers, FP32) against the oracle."""
    from lfric_core_amd import api
    from lfric_core_amd.configs import precision_config
    print("case C96L70 W2 mass fp32", flush=True)
    cfg = precision_config(96, 70, with_div_grad=False)
    M = cfg.mass.astype(np.float32)
    x = cfg.b.astype(np.float32)
    ref = oracle.invoke_coloured(0, cfg.colouring, 70, np.zeros(cfg.w2.undf, np.float32), x, M,
                                 cfg.w2.dofmap, cfg.w2.dofmap, nthreads=8)
    s2 = api.Space(ctx, cfg.w2)
    op = api.LMAOperator(_t(M, ctx), s2, s2, cfg.colouring)
    y = torch.empty(cfg.w2.undf, dtype=torch.float32, device=ctx.device)
    api.matrix_vector(y, _t(x, ctx), op, overwrite=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
