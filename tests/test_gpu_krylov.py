"""The generic iterative solvers (lfr_iterative_solve) over LMA, block,
dense (W0 + W3) and caller-written operators with null, diagonal, column-solve
and caller-written preconditioners, against the oracle's restatements of
sci_iterative_solver_mod.f90 on the same inputs; n-field vector reductions;
mm_diagonal bit-exact against the oracle."""
import numpy as np
import pytest

import oracle
from oracle import solvers as osol
from lfric_core_amd.configs import normal_lma, precision_config, spd_lma
from lfric_core_amd.function_space import build_dofmap, colour_cells
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def prob(ctx):
    from lfric_core_amd import api
    cfg = precision_config(4, 6)
    return cfg, api.Space(ctx, cfg.w2), api.Space(ctx, cfg.w3)


def _t(ctx, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)


def _w2_apply(cfg, M):
    return lambda v: oracle.invoke_coloured(0, cfg.colouring, cfg.nlayers, np.zeros_like(v), v, M,
                                            cfg.w2.dofmap, cfg.w2.dofmap)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# --- single-field W2 mass: every method against the oracle -------------------

METHODS = [
    ("cg", "ConjugateGradient", lambda ap, x, b, pr: osol.cg_solve(ap, x, b, prec=pr, r_tol=1e-10, max_iter=200)),
    ("bicgstab", "BiCGStab", lambda ap, x, b, pr: osol.bicgstab_solve(ap, x, b, prec=pr, r_tol=1e-10, max_iter=200)),
    ("gmres", "GMRES", lambda ap, x, b, pr: osol.gmres_solve(ap, x, b, prec=pr, gcrk=8, r_tol=1e-10, max_iter=50)),
    ("fgmres", "FGMRES", lambda ap, x, b, pr: osol.fgmres_solve(ap, x, b, prec=pr, gcrk=8, r_tol=1e-10, max_iter=50)),
    ("gcr", "GCR", lambda ap, x, b, pr: osol.gcr_solve(ap, x, b, prec=pr, gcrk=8, r_tol=1e-10, max_iter=50)),
]


@pytest.mark.parametrize("name,cls,orc", METHODS, ids=[m[0] for m in METHODS])
@pytest.mark.parametrize("precname", ["null", "diag"])
def test_solver_w2_mass_fp64_matches_oracle(ctx, prob, name, cls, orc, precname):
    from lfric_core_amd import api, solvers
    cfg, w2, _ = prob
    print(f"case {name} {precname} C4L6 fp64", flush=True)
    M = cfg.mass
    diag = oracle.mm_diagonal(np.zeros(cfg.w2.undf), M, cfg.w2.dofmap, cfg.nlayers, cfg.colouring)
    pr = (lambda v: v / diag) if precname == "diag" else None
    x0 = Stream(50).uniform(cfg.w2.undf) * 0.1
    xo = x0.copy()
    out = orc(_w2_apply(cfg, M), xo, cfg.b, pr)
    op = api.LMAOperator(_t(ctx, M), w2, w2, cfg.colouring)
    lin = solvers.MassMatrixOperator(op)          # the generic path for every method
    prec = solvers.DiagonalPreconditioner(_t(ctx, diag)) if precname == "diag" else None
    kw = dict(gcrk=8, max_iter=50) if name in ("gmres", "fgmres", "gcr") else dict(max_iter=200)
    s = getattr(solvers, cls)(lin, prec, r_tol=1e-10, **kw)
    x = _t(ctx, x0)
    st = s.apply(x, _t(ctx, cfg.b))
    assert st.converged
    assert st.iterations == out[0], (st.iterations, out[0])
    assert _rel(x.cpu().numpy(), xo) <= 1e-10


def test_fast_and_generic_cg_agree(ctx, prob):
    """CG through lfr_cg and through lfr_iterative_solve: same iteration count
    and solutions within rounding of the dot order."""
    from lfric_core_amd import api, solvers
    cfg, w2, _ = prob
    op = api.LMAOperator(_t(ctx, cfg.mass), w2, w2, cfg.colouring)
    b = _t(ctx, cfg.b)
    x1 = torch.zeros_like(b)
    s1 = solvers.ConjugateGradient(op, r_tol=1e-10, max_iter=200).apply(x1, b)
    x2 = torch.zeros_like(b)
    s2 = solvers.ConjugateGradient(solvers.MassMatrixOperator(op), r_tol=1e-10, max_iter=200).apply(x2, b)
    assert s1.iterations == s2.iterations and s1.converged and s2.converged
    assert _rel(x2.cpu().numpy(), x1.cpu().numpy()) <= 1e-12


def test_bicgstab_fp32_vectors(ctx, prob):
    """r_solver vectors: FP32 updates with FP64 scalars and inner products."""
    from lfric_core_amd import api, solvers
    cfg, w2, _ = prob
    M = cfg.mass.astype(np.float32)
    b = cfg.b.astype(np.float32)
    xo = np.zeros(cfg.w2.undf, np.float32)
    osol.bicgstab_solve(_w2_apply(cfg, M), xo, b, max_iter=12, monitor=False)
    op = api.LMAOperator(_t(ctx, M), w2, w2, cfg.colouring)
    x = torch.zeros(cfg.w2.undf, dtype=torch.float32, device=ctx.device)
    st = solvers.BiCGStab(op, max_iter=12, monitor_convergence=False).apply(x, _t(ctx, b))
    assert st.iterations == 12
    assert _rel(x.cpu().numpy(), xo) <= 2e-5


# --- W3 operator written by the caller, column-solve preconditioner -------------

def _w3_problem(ctx, cfg, w2, w3):
    """A = D G + c I on W3 (D: div W2 -> W3, G: grad W3 -> W2) as a callback
    operator; P = columnwise_op_appinv of a tridiagonal column operator."""
    from lfric_core_amd import api
    c = 30.0
    D, G = cfg.div, cfg.grad
    Dop = api.LMAOperator(_t(ctx, D), w3, w2, None)
    Gop = api.LMAOperator(_t(ctx, G), w2, w3, cfg.colouring)
    tmp = torch.empty(cfg.w2.undf, dtype=torch.float64, device=ctx.device)

    def apply_gpu(x, y):
        api.matrix_vector(tmp, x, Gop, overwrite=True)
        api.dg_matrix_vector(y, tmp, Dop)
        api.axpy(w3.ctx, c, x, y)

    def apply_orc(v):
        t = oracle.invoke_coloured(0, cfg.colouring, cfg.nlayers, np.zeros(cfg.w2.undf), v, G,
                                   cfg.w2.dofmap, cfg.w3.dofmap)
        y = oracle.invoke_dg(0, cfg.nlayers, np.zeros(cfg.w3.undf), t, D, cfg.w3.dofmap, cfg.w2.dofmap)
        return y + c * v

    nl, nc = cfg.nlayers, cfg.mesh.ncells
    st = Stream(51)
    cm = np.empty((nc, nl, 3))
    cm[:, :, 1] = c + st.uniform(nc * nl, 1.0, 3.0).reshape(nc, nl)
    cm[:, :, 0] = -1.0 + 0.1 * st.uniform(nc * nl).reshape(nc, nl)
    cm[:, :, 2] = -1.0 + 0.1 * st.uniform(nc * nl).reshape(nc, nl)
    cm[:, 0, 0] = 0.0
    cm[:, -1, 2] = 0.0
    ind = np.arange(1, nl + 1, dtype=np.int32)
    cma = api.CMAOperator(_t(ctx, cm), w3, ind)
    prec_orc = lambda v: oracle.appinv(np.zeros_like(v), v, cm, cfg.w3.dofmap, ind)   # noqa: E731
    return apply_gpu, apply_orc, cma, prec_orc


@pytest.mark.parametrize("name", ["bicgstab", "fgmres", "gcr", "gmres", "cg_jacobi"])
def test_callback_operator_with_column_preconditioner(ctx, prob, name):
    from lfric_core_amd import solvers
    cfg, w2, w3 = prob
    print(f"case callback W3 operator + appinv {name}", flush=True)
    apply_gpu, apply_orc, cma, prec_orc = _w3_problem(ctx, cfg, w2, w3)
    b = Stream(52).uniform(cfg.w3.undf)
    op = solvers.CallbackOperator(ctx, [cfg.w3.undf], torch.float64, apply_gpu)
    prec = solvers.ColumnwisePreconditioner(cma)
    xo = np.zeros(cfg.w3.undf)
    x = torch.zeros(cfg.w3.undf, dtype=torch.float64, device=ctx.device)
    if name == "bicgstab":
        out = osol.bicgstab_solve(apply_orc, xo, b, prec=prec_orc, r_tol=1e-10, max_iter=100)
        st = solvers.BiCGStab(op, prec, r_tol=1e-10, max_iter=100).apply(x, _t(ctx, b))
    elif name == "fgmres":
        out = osol.fgmres_solve(apply_orc, xo, b, prec=prec_orc, gcrk=5, r_tol=1e-10, max_iter=40)
        st = solvers.FGMRES(op, prec, gcrk=5, r_tol=1e-10, max_iter=40).apply(x, _t(ctx, b))
    elif name == "gcr":
        out = osol.gcr_solve(apply_orc, xo, b, prec=prec_orc, gcrk=5, r_tol=1e-10, max_iter=40)
        st = solvers.GCR(op, prec, gcrk=5, r_tol=1e-10, max_iter=40).apply(x, _t(ctx, b))
    elif name == "gmres":
        out = osol.gmres_solve(apply_orc, xo, b, prec=prec_orc, gcrk=5, r_tol=1e-10, max_iter=40)
        st = solvers.GMRES(op, prec, gcrk=5, r_tol=1e-10, max_iter=40).apply(x, _t(ctx, b))
    else:
        out = osol.jacobi_solve(apply_orc, xo, b, prec=prec_orc, r_tol=1e-9, max_iter=300, rho_relax=0.9)
        st = solvers.Jacobi(op, prec, r_tol=1e-9, max_iter=300, rho_relax=0.9).apply(x, _t(ctx, b))
    assert st.converged and out[1]
    assert st.iterations == out[0]
    assert _rel(x.cpu().numpy(), xo) <= 1e-9
    np.testing.assert_allclose(apply_orc(xo), b, atol=1e-7 * np.abs(b).max())


def test_solve_bytes_closed_form(ctx, prob):
    """lfr_iterative_solve_bytes: Jacobi with monitoring off makes, per
    iteration, one apply (matrix, x read, y written), two axpys (3 passes
    each) and one null-preconditioner copy (2 passes): 3 iterations count
    3 (M + 2 undf s + 8 undf s) bytes exactly.  A caller-written operator
    makes the count unknown."""
    from lfric_core_amd import api, solvers
    cfg, w2, w3 = prob
    print("case solve bytes Jacobi W2 C4L6 fp64", flush=True)
    op = api.LMAOperator(_t(ctx, cfg.mass), w2, w2, cfg.colouring)
    s = solvers.Jacobi(solvers.MassMatrixOperator(op), None, max_iter=3, monitor_convergence=False,
                       rho_relax=0.5)
    x = torch.zeros(cfg.w2.undf, dtype=torch.float64, device=ctx.device)
    s.apply(x, _t(ctx, cfg.b))
    u, n3 = cfg.w2.undf, cfg.mesh.ncells * cfg.nlayers
    assert s.bytes == 3 * (n3 * 36 * 8 + 2 * u * 8 + 8 * u * 8)
    apply_gpu, _, _, _ = _w3_problem(ctx, cfg, w2, w3)
    cb = solvers.CallbackOperator(ctx, [cfg.w3.undf], torch.float64, apply_gpu)
    s2 = solvers.Jacobi(cb, None, max_iter=2, monitor_convergence=False)
    s2.apply(torch.zeros(cfg.w3.undf, dtype=torch.float64, device=ctx.device),
             torch.ones(cfg.w3.undf, dtype=torch.float64, device=ctx.device))
    assert s2.bytes is None


def test_chebyshev_and_precondition_only(ctx, prob):
    from lfric_core_amd import api, solvers
    cfg, w2, _ = prob
    M = cfg.mass
    diag = oracle.mm_diagonal(np.zeros(cfg.w2.undf), M, cfg.w2.dofmap, cfg.nlayers, cfg.colouring)
    op = api.LMAOperator(_t(ctx, M), w2, w2, cfg.colouring)
    prec = solvers.DiagonalPreconditioner(_t(ctx, diag))
    xo = np.zeros(cfg.w2.undf)
    i0, f0 = osol.chebyshev_solve(_w2_apply(cfg, M), xo, cfg.b, prec=lambda v: v / diag, max_iter=15,
                                  lmin=0.2, lmax=2.5)
    x = torch.full((cfg.w2.undf,), 3.0, dtype=torch.float64, device=ctx.device)
    st = solvers.Chebyshev(op, prec, max_iter=15, lmin=0.2, lmax=2.5).apply(x, _t(ctx, cfg.b))
    assert _rel(x.cpu().numpy(), xo) <= 1e-12
    assert abs(st.r_nrm - f0) <= 1e-10 * f0 and abs(st.r_nrm0 - i0) <= 1e-12 * i0
    xp = np.zeros(cfg.w2.undf)
    osol.precondition_only_solve(_w2_apply(cfg, M), xp, cfg.b, prec=lambda v: v / diag)
    x = torch.zeros(cfg.w2.undf, dtype=torch.float64, device=ctx.device)
    solvers.PreconditionOnly(op, prec).apply(x, _t(ctx, cfg.b))
    np.testing.assert_array_equal(x.cpu().numpy(), xp)


# --- dense_operator_type: W0 + W3 two-field vectors -----------------------------

@pytest.fixture(scope="module")
def dense(ctx):
    from lfric_core_amd import api
    m = cubed_sphere(4)
    nl = 5
    w0 = build_dofmap(m, "W0", nl)
    w3 = build_dofmap(m, "W3", nl)
    col = colour_cells(m)
    st = Stream(53)
    n3 = m.ncells * nl
    M0 = spd_lma(st, n3, 8, 8.0)
    M3 = spd_lma(st, n3, 1, 1.0)
    s0, s3 = api.Space(ctx, w0), api.Space(ctx, w3)
    return dict(m=m, nl=nl, w0=w0, w3=w3, col=col, M0=M0, M3=M3, s0=s0, s3=s3)


def _dense_apply(d):
    def ap(v):
        n0 = d["w0"].undf
        y0 = oracle.invoke_coloured(0, d["col"], d["nl"], np.zeros(n0), v[:n0].copy(), d["M0"],
                                    d["w0"].dofmap, d["w0"].dofmap)
        y3 = oracle.invoke_dg(1, d["nl"], np.zeros(d["w3"].undf), v[n0:].copy(), d["M3"],
                              d["w3"].dofmap, d["w3"].dofmap)
        return np.concatenate([y0, y3])
    return ap


def test_dense_operator_apply_bitexact_and_diagonal(ctx, dense):
    from lfric_core_amd import api, solvers
    d = dense
    op0 = api.LMAOperator(_t(ctx, d["M0"]), d["s0"], d["s0"], d["col"])
    op3 = api.LMAOperator(_t(ctx, d["M3"]), d["s3"], d["s3"], None)
    A = solvers.DenseOperator(op0, op3)
    fields = [d["w0"].undf, d["w3"].undf]
    x = Stream(54).uniform(sum(fields))
    X = solvers.FieldVector(ctx, fields, torch.float64, _t(ctx, x))
    Y = X.duplicate()
    A.apply(X, Y)
    np.testing.assert_array_equal(Y.data.cpu().numpy(), _dense_apply(d)(x))
    Dg = A.extract_diagonal(X.duplicate())
    d0 = oracle.mm_diagonal(np.zeros(fields[0]), d["M0"], d["w0"].dofmap, d["nl"], d["col"])
    d3 = oracle.mm_diagonal(np.zeros(fields[1]), d["M3"], d["w3"].dofmap, d["nl"], None)
    np.testing.assert_array_equal(Dg.data.cpu().numpy(), np.concatenate([d0, d3]))
    assert X.vector_size() == 2


def test_field_vector_reductions(ctx, dense):
    from lfric_core_amd import api, solvers
    fields = [dense["w0"].undf, dense["w3"].undf]
    st = Stream(55)
    x, y = st.uniform(sum(fields)), st.uniform(sum(fields))
    X = solvers.FieldVector(ctx, fields, torch.float64, _t(ctx, x))
    Y = solvers.FieldVector(ctx, fields, torch.float64, _t(ctx, y))
    # the n-field dot is the field dots added in order (each the single-field reduction)
    a = api.dot(ctx, X.field(0), Y.field(0))
    b = api.dot(ctx, X.field(1), Y.field(1))
    assert X.dot(Y) == (0.0 + a) + b
    assert abs(X.dot(Y) - osol._dot(x, y, fields)) <= 1e-13 * abs(X.dot(Y))
    assert abs(X.norm() - osol._norm(x, fields)) <= 1e-14 * X.norm()
    assert abs(X.field_norm(2) - np.linalg.norm(x[fields[0]:])) <= 1e-14 * X.field_norm(2)
    assert X.field_norm(1) == np.sqrt(api.dot(ctx, X.field(0), X.field(0)))


@pytest.mark.parametrize("name", ["cg", "bicgstab", "gcr"])
def test_dense_operator_solves_match_oracle(ctx, dense, name):
    from lfric_core_amd import api, solvers
    d = dense
    print(f"case dense W0+W3 {name}", flush=True)
    op0 = api.LMAOperator(_t(ctx, d["M0"]), d["s0"], d["s0"], d["col"])
    op3 = api.LMAOperator(_t(ctx, d["M3"]), d["s3"], d["s3"], None)
    A = solvers.DenseOperator(op0, op3)
    fields = [d["w0"].undf, d["w3"].undf]
    b = Stream(56).uniform(sum(fields))
    B = solvers.FieldVector(ctx, fields, torch.float64, _t(ctx, b))
    Dg = A.extract_diagonal(B.duplicate())
    dg = Dg.data.cpu().numpy()
    X = B.duplicate()
    xo = np.zeros(sum(fields))
    ap = _dense_apply(d)
    pr = lambda v: v / dg                              # noqa: E731
    prec = solvers.DiagonalPreconditioner(Dg)
    if name == "cg":
        out = osol.cg_solve(ap, xo, b, prec=pr, r_tol=1e-11, max_iter=300, fields=fields)
        st = solvers.ConjugateGradient(A, prec, r_tol=1e-11, max_iter=300).apply(X, B)
    elif name == "bicgstab":
        out = osol.bicgstab_solve(ap, xo, b, prec=pr, r_tol=1e-11, max_iter=300, fields=fields)
        st = solvers.BiCGStab(A, prec, r_tol=1e-11, max_iter=300).apply(X, B)
    else:
        out = osol.gcr_solve(ap, xo, b, prec=pr, gcrk=10, r_tol=1e-11, max_iter=30, fields=fields)
        st = solvers.GCR(A, prec, gcrk=10, r_tol=1e-11, max_iter=30).apply(X, B)
    assert st.converged and out[1]
    assert st.iterations == out[0]
    assert _rel(X.data.cpu().numpy(), xo) <= 1e-10


def test_block_operator_with_transpose_entry(ctx, prob):
    """[[M, G], [D, c I]]-shaped saddle-like block operator: the transpose entry
    (D^T into W2) and the grad entry against the oracle's invokes, bit-exact."""
    from lfric_core_amd import api, solvers
    cfg, w2, w3 = prob
    D, M = cfg.div, cfg.mass
    Mop = api.LMAOperator(_t(ctx, M), w2, w2, cfg.colouring)
    Dop_t = api.LMAOperator(_t(ctx, D), w3, w2, cfg.colouring)   # D^T: W3 -> W2 with W2's colouring
    Dop = api.LMAOperator(_t(ctx, D), w3, w2, None)
    A = solvers.BlockOperator([w2, w3], [(0, 0, Mop), (0, 1, Dop_t, True), (1, 0, Dop)])
    n2, n3 = cfg.w2.undf, cfg.w3.undf
    x = Stream(57).uniform(n2 + n3)
    y = torch.empty(n2 + n3, dtype=torch.float64, device=ctx.device)
    A.apply(_t(ctx, x), y)
    y0 = oracle.invoke_coloured(0, cfg.colouring, cfg.nlayers, np.zeros(n2), x[:n2].copy(), M,
                                cfg.w2.dofmap, cfg.w2.dofmap)
    y0 = oracle.invoke_coloured(1, cfg.colouring, cfg.nlayers, y0, x[n2:].copy(), D, cfg.w2.dofmap, cfg.w3.dofmap)
    y1 = oracle.invoke_dg(1, cfg.nlayers, np.zeros(n3), x[:n2].copy(), D, cfg.w3.dofmap, cfg.w2.dofmap)
    np.testing.assert_array_equal(y.cpu().numpy(), np.concatenate([y0, y1]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mm_diagonal_gpu_bitexact(ctx, dtype):
    from lfric_core_amd import api
    ctx.sync_stream()          # the raw calls below queue on PyTorch's current stream
    m = cubed_sphere(5)
    nl = 7
    col = colour_cells(m)
    st = Stream(58)
    for name in ("W0", "W2", "WTHETA", "W3"):
        fs = build_dofmap(m, name, nl)
        sp = api.Space(ctx, fs)
        M = normal_lma(st, m.ncells * nl, fs.ndf, fs.ndf).astype(dtype)
        d0 = st.uniform(fs.undf).astype(dtype)
        c = None if fs.is_discontinuous else col
        ref = oracle.mm_diagonal(d0.copy(), M, fs.dofmap, nl, c)
        g = _t(ctx, d0.copy())
        plan = sp.plan(col, 0) if c is not None else None
        from lfric_core_amd._ffi import check, lib
        import ctypes as C
        check(lib().lfr_mm_diagonal(ctx.h, 4 if dtype == np.float32 else 8, C.c_void_p(g.data_ptr()),
                                    C.c_void_p(_t(ctx, M).data_ptr()), sp.h, plan.h if plan else None), ctx.h)
        np.testing.assert_array_equal(g.cpu().numpy(), ref, err_msg=name)


def test_error_paths(ctx, prob):
    from lfric_core_amd import api, solvers
    from lfric_core_amd._ffi import LfrError
    cfg, w2, w3 = prob

    def bad(x, y):
        raise RuntimeError("boom")
    op = solvers.CallbackOperator(ctx, [cfg.w3.undf], torch.float64, bad)
    b = torch.ones(cfg.w3.undf, dtype=torch.float64, device=ctx.device)
    with pytest.raises(LfrError, match="callback"):
        solvers.BiCGStab(op).apply(torch.zeros_like(b), b)
    good = solvers.CallbackOperator(ctx, [cfg.w3.undf], torch.float64, lambda x, y: y.copy_(x))
    with pytest.raises(LfrError):
        solvers.GCR(good, solvers.DiagonalPreconditioner(torch.ones(5, dtype=torch.float64,
                                                                    device=ctx.device))).apply(torch.zeros_like(b), b)
    # identity operator: CG converges in one iteration with x = b
    x = torch.zeros_like(b)
    st = solvers.ConjugateGradient(good, r_tol=1e-12, max_iter=5).apply(x, b)
    assert st.converged and st.iterations == 1
    np.testing.assert_array_equal(x.cpu().numpy(), b.cpu().numpy())
    # GCR starts from dx = P^{-1} b, exact here: its next direction is zero and
    # the reference divides by its norm (NaN residual, reported as an error)
    with pytest.raises(LfrError, match="NaN"):
        solvers.GCR(good, gcrk=3, r_tol=1e-12, max_iter=2).apply(torch.zeros_like(b), b)
    # not converged with fail_on_non_converged
    M = cfg.mass
    opm = api.LMAOperator(_t(ctx, M), w2, w2, cfg.colouring)
    with pytest.raises(LfrError, match="NOT converged"):
        solvers.BiCGStab(opm, r_tol=1e-14, max_iter=2, fail_on_non_converged=True).apply(
            torch.zeros(cfg.w2.undf, dtype=torch.float64, device=ctx.device), _t(ctx, cfg.b))
