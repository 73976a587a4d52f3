"""16-field MFMA apply (lfr_matrix_vector_multi) against the oracle's FP64
single-field loops, one field at a time.

Tolerances (norm-wise relative error per field, against FP64 on FP64 data):
FP32 MFMA 2e-6; BF16-input/FP32-accumulate 2e-2 (BF16 keeps 8 mantissa bits).
The MFMA summation order differs from the single-field kernels, so these
are tolerances, not bit-exactness; repeated applies are bitwise identical."""
import numpy as np
import pytest

import oracle
from lfric_core_amd import function_space as fsm
from lfric_core_amd.configs import normal_lma, spd_lma
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = {4: 2e-6, 2: 2e-2}
CASES = [  # (to space, k, from space, k, matrix kind)
    ("W2", 1, "W2", 1, "spd"),
    ("W0", 1, "W0", 1, "spd"),
    ("W3", 1, "W3", 1, "spd"),
    ("W3", 1, "W2", 1, "normal"),
    ("W0", 0, "W0", 0, "spd"),
    ("W2", 0, "W2", 0, "spd"),
]


def _problem(to, kt, fr, kf, kind, nl, edge=4, seed=3):
    mesh = cubed_sphere(edge)
    fs_to = fsm.build_dofmap(mesh, to, nl, k=kt)
    fs_fr = fs_to if (fr, kf) == (to, kt) else fsm.build_dofmap(mesh, fr, nl, k=kf)
    col = fsm.colour_cells(mesh)
    st = Stream(seed)
    n3 = mesh.ncells * nl
    if kind == "spd":
        M = spd_lma(st, n3, fs_to.ndf, float(fs_to.ndf), std=1.0 / np.sqrt(fs_to.ndf))
    else:
        M = normal_lma(st, n3, fs_to.ndf, fs_fr.ndf)
    X = st.uniform(fs_fr.undf * 16).reshape(fs_fr.undf, 16)
    return mesh, fs_to, fs_fr, col, M, X


def _oracle(fs_to, fs_fr, col, nl, M, X):
    Y = np.zeros((fs_to.undf, 16))
    for r in range(16):
        x = np.ascontiguousarray(X[:, r])
        if fs_to.is_discontinuous:
            y = oracle.invoke_dg(0, nl, np.zeros(fs_to.undf), x, M, fs_to.dofmap, fs_fr.dofmap)
        else:
            y = oracle.invoke_coloured(0, col, nl, np.zeros(fs_to.undf), x, M, fs_to.dofmap,
                                       fs_fr.dofmap)
        Y[:, r] = y
    return Y


@pytest.mark.parametrize("mprec", [4, 2])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}k{c[1]}-{c[2]}k{c[3]}")
def test_multi_apply_matches_fp64(ctx, case, mprec):
    from lfric_core_amd import api
    nl = 20                                     # one full 16-layer block and a partial one
    to, kt, fr, kf, kind = case
    print(f"multi {to}k{kt}<-{fr}k{kf} mprec={mprec}", flush=True)
    mesh, fs_to, fs_fr, col, M, X = _problem(to, kt, fr, kf, kind, nl)
    Y64 = _oracle(fs_to, fs_fr, col, nl, M, X)
    s_to = api.Space(ctx, fs_to)
    s_fr = s_to if fs_fr is fs_to else api.Space(ctx, fs_fr)
    op = api.LMAOperator(torch.from_numpy(M.astype(np.float32)).to(ctx.device), s_to, s_fr, col)
    mop = api.MultiOperator(op, mprec)
    xd = torch.from_numpy(X.astype(np.float32)).to(ctx.device)
    y = torch.full((fs_to.undf, 16), np.nan, dtype=torch.float32, device=ctx.device)
    mop.apply(y, xd, overwrite=True)
    Y = y.cpu().numpy().astype(np.float64)
    err = np.linalg.norm(Y - Y64, axis=0) / np.linalg.norm(Y64, axis=0)
    assert np.all(np.isfinite(Y)), "unwritten output dofs"
    assert err.max() < TOL[mprec], err
    # increment form adds onto lhs; repeated overwrite applies are bitwise identical
    y2 = y.clone()
    mop.apply(y2, xd, overwrite=False)
    np.testing.assert_allclose(y2.cpu().numpy(), 2 * Y, rtol=1e-6, atol=1e-6 * np.abs(Y).max())
    y3 = torch.zeros_like(y)
    mop.apply(y3, xd, overwrite=True)
    assert torch.equal(y3, y)
    mop.close()


def test_multi_info_and_errors(ctx):
    from lfric_core_amd import api
    from lfric_core_amd._ffi import LfrError
    mesh, fs_to, fs_fr, col, M, X = _problem("W2", 1, "W2", 1, "spd", 16, edge=2)
    s = api.Space(ctx, fs_to)
    op = api.LMAOperator(torch.from_numpy(M.astype(np.float32)).to(ctx.device), s, s, col)
    mop = api.MultiOperator(op, 4)
    info = mop.info()
    # owner-computes form: one launch over all cells
    assert info["slots"] == 36 and info["vertical_pairs"] == 4 and info["launches"] == 1
    with pytest.raises(LfrError):
        mop.apply(torch.zeros(fs_to.undf, 8, device=ctx.device),
                  torch.zeros(fs_to.undf, 16, device=ctx.device))
    mop.close()


@pytest.mark.parametrize("mprec", [4, 2])
@pytest.mark.parametrize("edge,nl", [(4, 20), (6, 70), (2, 16)])
def test_multi_owner_form_matches_colour_form(ctx, monkeypatch, mprec, edge, nl):
    """The owner-computes form (one launch, each output written once) and the
    colour-by-colour form (LFR_MULTI_OWNER=0) give the same bits in the
    overwrite form: each cell's contribution is the same MFMA sequence and a
    shared dof is the sum of exactly two contributions, a + b == b + a."""
    from lfric_core_amd import api
    print(f"multi owner vs colour W2k1 C{edge}L{nl} mprec={mprec}", flush=True)
    mesh, fs, _, col, M, X = _problem("W2", 1, "W2", 1, "spd", nl, edge=edge, seed=11)
    s = api.Space(ctx, fs)
    op = api.LMAOperator(torch.from_numpy(M.astype(np.float32)).to(ctx.device), s, s, col)
    xd = torch.from_numpy(X.astype(np.float32)).to(ctx.device)
    own = api.MultiOperator(op, mprec)
    monkeypatch.setenv("LFR_MULTI_OWNER", "0")
    clr = api.MultiOperator(op, mprec)
    assert own.info()["launches"] == 1 and clr.info()["launches"] == col.ncolours
    ya = torch.full((fs.undf, 16), np.nan, dtype=torch.float32, device=ctx.device)
    yb = torch.full_like(ya, np.nan)
    own.apply(ya, xd, overwrite=True)
    clr.apply(yb, xd, overwrite=True)
    assert torch.isfinite(ya).all()
    assert torch.equal(ya, yb)
    # increment form: y + (a + b) against (y + a) + b, equal to rounding
    za, zb = ya.clone(), ya.clone()
    own.apply(za, xd, overwrite=False)
    clr.apply(zb, xd, overwrite=False)
    torch.testing.assert_close(za, zb, rtol=1e-6, atol=1e-6 * float(ya.abs().max()))
    # the owner increment adds y to the very sum the overwrite form stores
    # (y loaded ahead of the stores, layers past the top of a partial last
    # block clamped): with y = that sum the result is y + y exactly
    assert torch.equal(za, ya + ya)
    own.close()
    clr.close()


def test_refined_cg_multi_16_fields(ctx):
    """16 simultaneous refined solves to 1e-10; residuals re-measured with the
    oracle's FP64 coloured apply on the CPU."""
    from lfric_core_amd import api
    from lfric_core_amd.solvers import RefinedCGMulti
    nl = 20
    mesh, fs, _, col, M, X = _problem("W2", 1, "W2", 1, "spd", nl)
    s = api.Space(ctx, fs)
    M64 = torch.from_numpy(M).to(ctx.device)
    op64 = api.LMAOperator(M64, s, s, col)
    op32 = api.LMAOperator(M64.float().contiguous(), s, s, col)
    mop = api.MultiOperator(op32, 4)
    b = torch.from_numpy(X).to(ctx.device)
    x = torch.zeros_like(b)
    st = RefinedCGMulti(mop, op64, r_tol=1e-10).apply(x, b)
    assert st.converged and st.r_nrm <= 1e-10 and st.outer >= 1
    xs = x.cpu().numpy()
    for r in range(16):
        Ax = oracle.invoke_coloured(0, col, nl, np.zeros(fs.undf), np.ascontiguousarray(xs[:, r]), M,
                                    fs.dofmap, fs.dofmap)
        rel = np.linalg.norm(X[:, r] - Ax) / np.linalg.norm(X[:, r])
        assert rel <= 1.5e-10, (r, rel)
    mop.close()
