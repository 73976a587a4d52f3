"""Packed operators (lfr_lma_pack): applies through the gather-order copy of
an LMA matrix give the same bits as the oracle's coloured loops, for the
forward and transposed kernels, GH_INC and overwrite, FP32 and FP64; the
solvers give the same iterates with a packed operator."""

import numpy as np
import pytest

import oracle
from lfric_core_amd.function_space import build_dofmap, colour_cells
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream
from lfric_core_amd.configs import normal_lma

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
    (5, 7, 0, "W2", "W2"),
    (6, 10, 0, "W0", "W0"),
    (6, 10, 0, "W2", "W3"),
    (3, 4, 1, "W0", "W3"),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl,k,to,frm", CASES)
def test_packed_matrix_vector_bitexact(ctx, dtype, nc, nl, k, to, frm):
    from lfric_core_amd import api
    print(f"case packed matrix_vector nc={nc} nl={nl} k={k} {to}<-{frm} {dtype.__name__}", flush=True)
    m, col, fs, sp = _setup(ctx, nc, nl, k, {to, frm})
    st = Stream(21)
    n3 = m.ncells * nl
    M = normal_lma(st, n3, fs[to].ndf, fs[frm].ndf).astype(dtype)
    x = st.uniform(fs[frm].undf).astype(dtype)
    y0 = st.uniform(fs[to].undf).astype(dtype)
    ref_inc = oracle.invoke_coloured(0, col, nl, y0.copy(), x, M, fs[to].dofmap, fs[frm].dofmap)
    ref_ovw = oracle.invoke_coloured(0, col, nl, np.zeros_like(y0), x, M, fs[to].dofmap, fs[frm].dofmap)
    Md = _t(M, ctx)
    op = api.LMAOperator(Md, sp[to], sp[frm], col)
    pk = api.PackedOperator(op, kind=0)
    # the pack owns its copy: clobbering the source matrix changes nothing
    Md.fill_(float("nan"))
    assert pk.nbytes >= M.nbytes
    y = _t(y0, ctx)
    pk.apply(y, _t(x, ctx))
    z = torch.full_like(y, float("nan"))
    pk.apply(z, _t(x, ctx), overwrite=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref_inc)
    np.testing.assert_array_equal(z.cpu().numpy(), ref_ovw)
    pk.close()


TCASES = [
    (4, 5, 0, "W3", "W2"),   # D^T: adjoint of div, output continuous W2, x on W3
    (6, 10, 0, "W2", "W2"),
    (6, 6, 0, "W0", "W0"),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl,k,to,frm", TCASES)
def test_packed_transpose_bitexact(ctx, dtype, nc, nl, k, to, frm):
    from lfric_core_amd import api
    print(f"case packed transpose nc={nc} nl={nl} k={k} {frm}<-{to}^T {dtype.__name__}", flush=True)
    m, col, fs, sp = _setup(ctx, nc, nl, k, {to, frm})
    st = Stream(22)
    n3 = m.ncells * nl
    M = normal_lma(st, n3, fs[to].ndf, fs[frm].ndf).astype(dtype)
    xt = st.uniform(fs[to].undf).astype(dtype)
    l0 = st.uniform(fs[frm].undf).astype(dtype)
    ref = oracle.invoke_coloured(1, col, nl, l0.copy(), xt, M, fs[frm].dofmap, fs[to].dofmap)
    op = api.LMAOperator(_t(M, ctx), sp[to], sp[frm], col)
    pk = api.PackedOperator(op, kind=1)
    lhs = _t(l0, ctx)
    pk.apply(lhs, _t(xt, ctx))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(lhs.cpu().numpy(), ref)


def test_packed_unsupported_and_bad_args(ctx):
    """36 input dofs per cell (W2, k=1) has no packed kernel: a clear error."""
    from lfric_core_amd import api
    from lfric_core_amd._ffi import LfrError
    print("case packed errors", flush=True)
    m, col, fs, sp = _setup(ctx, 2, 3, 1, {"W2"})
    n3 = m.ncells * 3
    M = torch.zeros((fs["W2"].ndf, fs["W2"].ndf, n3), dtype=torch.float32, device=ctx.device)
    with pytest.raises(LfrError) as e:
        api.PackedOperator(api.LMAOperator(M, sp["W2"], sp["W2"], col))
    assert e.value.code == 1 and "packed" in str(e.value)
    m, col, fs, sp = _setup(ctx, 2, 3, 0, {"W0"})
    M0 = torch.zeros((fs["W0"].ndf, fs["W0"].ndf, n3), dtype=torch.float64, device=ctx.device)
    pk = api.PackedOperator(api.LMAOperator(M0, sp["W0"], sp["W0"], col))
    with pytest.raises(LfrError):  # wrong dtype of x
        pk.apply(torch.zeros(fs["W0"].undf, dtype=torch.float64, device=ctx.device),
                 torch.zeros(fs["W0"].undf, dtype=torch.float32, device=ctx.device))


def test_packed_w2_c96l70_fp32_bitexact(ctx):
    """Full-size W2 mass apply (C96, 70 layers, FP32) through the pack."""
    from lfric_core_amd import api
    from lfric_core_amd.configs import precision_config
    print("case packed C96L70 W2 mass fp32", flush=True)
    cfg = precision_config(96, 70, with_div_grad=False)
    M = cfg.mass.astype(np.float32)
    x = cfg.b.astype(np.float32)
    ref = oracle.invoke_coloured(0, cfg.colouring, 70, np.zeros(cfg.w2.undf, np.float32), x, M,
                                 cfg.w2.dofmap, cfg.w2.dofmap, nthreads=8)
    s2 = api.Space(ctx, cfg.w2)
    pk = api.PackedOperator(api.LMAOperator(_t(M, ctx), s2, s2, cfg.colouring))
    y = torch.empty(cfg.w2.undf, dtype=torch.float32, device=ctx.device)
    pk.apply(y, _t(x, ctx), overwrite=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)


@pytest.fixture(scope="module")
def prob(ctx):
    from lfric_core_amd import api
    from lfric_core_amd.configs import precision_config
    cfg = precision_config(8, 10, with_div_grad=False)
    return cfg, api.Space(ctx, cfg.w2)


def _op(ctx, cfg, w2, dt, packed):
    from lfric_core_amd import api
    op = api.LMAOperator(_t(cfg.mass.astype(dt), ctx), w2, w2, cfg.colouring)
    if packed:
        op.packed = api.PackedOperator(op)
    return op


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cg_packed_identical(ctx, prob, dtype):
    """CG through the packed operator: the same iterate bits and counts."""
    from lfric_core_amd.solvers import ConjugateGradient
    cfg, w2 = prob
    print(f"case cg packed {dtype.__name__}", flush=True)
    b = _t(cfg.b.astype(dtype), ctx)
    xs = []
    for packed in (False, True):
        x = torch.zeros_like(b)
        st = ConjugateGradient(_op(ctx, cfg, w2, dtype, packed), r_tol=1e-6, max_iter=60).apply(x, b)
        xs.append((x.cpu().numpy(), st.iterations, st.r_nrm))
    np.testing.assert_array_equal(xs[0][0], xs[1][0])
    assert xs[0][1] == xs[1][1] and xs[0][2] == xs[1][2]


def test_refined_cg_packed_identical(ctx, prob):
    from lfric_core_amd.solvers import RefinedCG
    cfg, w2 = prob
    print("case refined cg packed", flush=True)
    b = _t(cfg.b, ctx)
    out = []
    for packed in (False, True):
        x = torch.zeros_like(b)
        st = RefinedCG(_op(ctx, cfg, w2, np.float32, packed), _op(ctx, cfg, w2, np.float64, packed),
                       r_tol=1e-10).apply(x, b)
        assert st.converged
        out.append((x.cpu().numpy(), st.inner_total, st.outer))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


def test_refined_inner_solve_after_gmres(ctx, prob):
    """The inner FP32 CG of the refinement sums exactly the partials its own
    apply wrote, whatever an earlier solver left in the shared workspace:
    after a GMRES run, one correction equals a standalone FP32 CG solve."""
    from lfric_core_amd._ffi import LfrError
    from lfric_core_amd.solvers import GMRES, ConjugateGradient, RefinedCG
    cfg, w2 = prob
    print("case refined inner solve after gmres", flush=True)
    op64 = _op(ctx, cfg, w2, np.float64, False)
    op32 = _op(ctx, cfg, w2, np.float32, False)
    b = _t(cfg.b, ctx)
    GMRES(op64, gcrk=12, r_tol=1e-12, max_iter=3).apply(torch.zeros_like(b), b)
    x = torch.zeros_like(b)
    with pytest.raises(LfrError):  # one correction does not reach 1e-14
        RefinedCG(op32, op64, r_tol=1e-14, max_outer=1, inner_r_tol=1e-4, inner_max_iter=40).apply(x, b)
    x32 = torch.zeros(cfg.w2.undf, dtype=torch.float32, device=ctx.device)
    ConjugateGradient(op32, r_tol=1e-4, max_iter=40).apply(x32, b.float(), x_is_zero=True)
    np.testing.assert_array_equal(x.cpu().numpy(), x32.cpu().numpy().astype(np.float64))


def test_cg_graph_cache_after_repack_and_new_plan(ctx):
    """CG on a non-default stream (the iteration graphs are captured and
    cached): solve with packed operator A, release the pack and the plan,
    pack a different matrix B through a new plan on the same vectors, solve
    again; both solves match the oracle.  A cached graph keyed on recycled
    host handles would replay A's freed buffers here."""
    from lfric_core_amd import api
    from lfric_core_amd.configs import precision_config, spd_lma
    from lfric_core_amd.solvers import ConjugateGradient
    from oracle.solvers import cg_solve
    cfg = precision_config(4, 6, with_div_grad=False)
    w2 = api.Space(ctx, cfg.w2)
    n3 = cfg.mesh.ncells * cfg.nlayers
    mats = [cfg.mass, spd_lma(Stream(59), n3, 6, 3.0)]
    b = _t(cfg.b, ctx)
    x = torch.zeros(cfg.w2.undf, dtype=torch.float64, device=ctx.device)
    s = torch.cuda.Stream()
    for trial, M in enumerate(mats):
        print(f"case cg graph cache trial {trial}", flush=True)
        xo = np.zeros(cfg.w2.undf)
        cg_solve(lambda v: oracle.invoke_coloured(0, cfg.colouring, cfg.nlayers, np.zeros_like(v), v, M,
                                                  cfg.w2.dofmap, cfg.w2.dofmap),
                 xo, cfg.b, max_iter=20, monitor=False)
        with torch.cuda.stream(s):
            op = api.LMAOperator(_t(M, ctx), w2, w2, cfg.colouring)
            pk = api.PackedOperator(op)
            op.packed = pk
            x.zero_()
            ConjugateGradient(op, max_iter=20, monitor_convergence=False).apply(x, b)
            got = x.cpu().numpy()
            pk.close()
            for p in list(w2._plans.values()):
                p.close()
            w2._plans.clear()
        s.synchronize()
        assert np.linalg.norm(got - xo) <= 1e-11 * np.linalg.norm(xo), trial
