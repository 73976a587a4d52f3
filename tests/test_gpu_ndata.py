"""Multidata applies (lfr_matrix_vector_ndata): every field of an
(undf, ndata) "ndata first" field equals, bit for bit, the single-field apply
of that field (which the LMA tests pin to the oracle), for matrix_vector and
transpose_matrix_vector, GH_INC and overwrite, FP32 and FP64."""

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
    # (nc, nl, k, to, from, ndata, transpose)
    (3, 4, 1, "W2", "W2", 16, False),   # 36 dofs per cell: plan-entry classes
    (4, 5, 0, "W2", "W2", 16, False),
    (4, 5, 0, "W2", "W2", 3, False),
    (6, 10, 0, "W0", "W0", 16, False),
    (4, 5, 0, "W3", "W2", 16, True),    # D^T onto W2
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl,k,to,frm,nd,tr", CASES)
def test_ndata_equals_single_field(ctx, dtype, nc, nl, k, to, frm, nd, tr):
    from lfric_core_amd import api
    print(f"case ndata nc={nc} nl={nl} k={k} {to}<-{frm} nd={nd} transpose={tr} {dtype.__name__}",
          flush=True)
    m, col, fs, sp = _setup(ctx, nc, nl, k, {to, frm})
    st = Stream(31)
    n3 = m.ncells * nl
    M = normal_lma(st, n3, fs[to].ndf, fs[frm].ndf).astype(dtype)
    op = api.LMAOperator(_t(M, ctx), sp[to], sp[frm], col)
    lsp, xsp = (frm, to) if tr else (to, frm)
    X = _t(st.uniform(fs[xsp].undf * nd).reshape(fs[xsp].undf, nd).astype(dtype), ctx)
    Y0 = _t(st.uniform(fs[lsp].undf * nd).reshape(fs[lsp].undf, nd).astype(dtype), ctx)
    Yi = Y0.clone()
    api.matrix_vector_ndata(Yi, X, op, transpose=tr)
    Yo = torch.full_like(Y0, float("nan"))
    api.matrix_vector_ndata(Yo, X, op, overwrite=True, transpose=tr)
    single = api.transpose_matrix_vector if tr else api.matrix_vector
    for r in range(nd):
        xr = X[:, r].contiguous()
        yi = Y0[:, r].contiguous()
        single(yi, xr, op)
        yo = torch.empty_like(yi)
        single(yo, xr, op, overwrite=True)
        assert torch.equal(Yi[:, r], yi), f"field {r} (increment)"
        assert torch.equal(Yo[:, r], yo), f"field {r} (overwrite)"
    if not tr and nd == 3:   # one direct oracle check
        ref = oracle.invoke_coloured(0, col, nl, np.zeros(fs[to].undf, dtype), X[:, 1].cpu().numpy(), M,
                                     fs[to].dofmap, fs[frm].dofmap)
        np.testing.assert_array_equal(Yo[:, 1].cpu().numpy(), ref)


def test_ndata_bad_args(ctx):
    from lfric_core_amd import api
    from lfric_core_amd._ffi import LfrError
    print("case ndata errors", flush=True)
    m, col, fs, sp = _setup(ctx, 2, 3, 0, {"W2"})
    M = torch.zeros((6, 6, m.ncells * 3), dtype=torch.float64, device=ctx.device)
    op = api.LMAOperator(M, sp["W2"], sp["W2"], col)
    u = fs["W2"].undf
    with pytest.raises(LfrError):   # ndata above 64
        api.matrix_vector_ndata(torch.zeros(u, 65, dtype=torch.float64, device=ctx.device),
                                torch.zeros(u, 65, dtype=torch.float64, device=ctx.device), op)
    with pytest.raises(LfrError):   # different ndata
        api.matrix_vector_ndata(torch.zeros(u, 4, dtype=torch.float64, device=ctx.device),
                                torch.zeros(u, 3, dtype=torch.float64, device=ctx.device), op)
