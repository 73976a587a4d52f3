"""GPU column solve (columnwise_op_appinv) against the oracle, bit-exact in
FP32 and FP64, plus the FP64-pivot mode against a banded direct solve."""
import numpy as np
import pytest

import oracle
from lfric_core_amd.function_space import build_dofmap
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _helmholtz_columns(st, ncell, nrow, h2=1.0):
    """diag = 2 + h^2 lambda, lambda ~ U[0.1, 10]; off-diagonals -1 +- 10% U."""
    cm = np.empty((ncell, nrow, 3))
    lam = st.uniform(ncell * nrow, 0.1, 10.0).reshape(ncell, nrow)
    cm[:, :, 1] = 2.0 + h2 * lam
    cm[:, :, 0] = -1.0 + 0.1 * st.uniform(ncell * nrow).reshape(ncell, nrow)
    cm[:, :, 2] = -1.0 + 0.1 * st.uniform(ncell * nrow).reshape(ncell, nrow)
    cm[:, 0, 0] = 0.0
    cm[:, -1, 2] = 0.0
    return cm


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl", [(6, 70), (3, 70), (4, 10), (3, 13)])
def test_appinv_bitexact(ctx, dtype, nc, nl):
    from lfric_core_amd import api
    print(f"case appinv C{nc} nl={nl} {dtype.__name__}", flush=True)
    m = cubed_sphere(nc)
    w3 = build_dofmap(m, "W3", nl)
    st = Stream(31)
    cm = _helmholtz_columns(st, m.ncells, nl).astype(dtype)
    x = st.uniform(w3.undf).astype(dtype)
    ind = np.arange(1, nl + 1, dtype=np.int32)
    ref = oracle.appinv(np.zeros_like(x), x, cm, w3.dofmap, ind)
    sp = api.Space(ctx, w3)
    op = api.CMAOperator(torch.from_numpy(cm).to(ctx.device), sp, ind)
    y = torch.full((w3.undf,), float("nan"), dtype=op.matrix.dtype, device=ctx.device)
    api.columnwise_op_appinv(y, torch.from_numpy(x).to(ctx.device), op)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)


def test_appinv_fp64_pivots(ctx):
    from scipy.linalg import solve_banded
    from lfric_core_amd import api
    m = cubed_sphere(4)
    nl = 70
    w3 = build_dofmap(m, "W3", nl)
    st = Stream(32)
    cm = _helmholtz_columns(st, m.ncells, nl, h2=1e-4).astype(np.float32)   # near-singular rows
    x = st.uniform(w3.undf).astype(np.float32)
    ind = np.arange(1, nl + 1, dtype=np.int32)
    ref64 = oracle.appinv(np.zeros(w3.undf), x.astype(np.float64), cm.astype(np.float64),
                          w3.dofmap, ind)
    sp = api.Space(ctx, w3)
    op = api.CMAOperator(torch.from_numpy(cm).to(ctx.device), sp, ind)
    y = torch.empty(w3.undf, dtype=torch.float32, device=ctx.device)
    api.columnwise_op_appinv(y, torch.from_numpy(x).to(ctx.device), op, fp64_pivots=True)
    got = y.cpu().numpy().astype(np.float64)
    # FP64 recurrence rounded once to FP32: within one FP32 ulp of the FP64 solve
    np.testing.assert_allclose(got, ref64.astype(np.float32).astype(np.float64), rtol=2e-7, atol=0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc", [3, 6])
def test_appinv_permuted_indirection(ctx, dtype, nc):
    """A non-identity indirection_dofmap (rows solved in a permuted dof order)
    at 70 rows, where the FP32 solve reads and writes through the indirection
    table in flat order; 54 and 216 columns leave partial workgroups."""
    from lfric_core_amd import api
    print(f"case appinv permuted indirection C{nc} nl=70 {dtype.__name__}", flush=True)
    m = cubed_sphere(nc)
    w3 = build_dofmap(m, "W3", 70)
    st = Stream(37)
    cm = _helmholtz_columns(st, m.ncells, 70).astype(dtype)
    x = st.uniform(w3.undf).astype(dtype)
    ind = (np.random.default_rng(5).permutation(70) + 1).astype(np.int32)
    ref = oracle.appinv(np.zeros_like(x), x, cm, w3.dofmap, ind)
    sp = api.Space(ctx, w3)
    op = api.CMAOperator(torch.from_numpy(cm).to(ctx.device), sp, ind)
    y = torch.full((w3.undf,), float("nan"), dtype=op.matrix.dtype, device=ctx.device)
    api.columnwise_op_appinv(y, torch.from_numpy(x).to(ctx.device), op)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
