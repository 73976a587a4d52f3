"""Vertical tridiagonal solves on the GPU (lfr_tri_solve, lfr_operator_tri_solve)
against the oracle, bit for bit (same reciprocal-then-multiply order)."""

import numpy as np
import pytest

import oracle
from lfric_core_amd.function_space import build_dofmap
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _t(a, ctx):
    return torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nc,nl", [(3, 10), (8, 70), (9, 70)])   # C9: a partial workgroup of columns
def test_tri_solve_bitexact(ctx, dtype, nc, nl):
    from lfric_core_amd import api
    print(f"case tri_solve nc={nc} nl={nl} {dtype.__name__}", flush=True)
    fs = build_dofmap(cubed_sphere(nc), "W3", nl)
    st = Stream(51)
    t0 = (4.0 + st.uniform(fs.undf)).astype(dtype)
    tp = (0.5 * st.uniform(fs.undf)).astype(dtype)
    tm = (0.5 * st.uniform(fs.undf)).astype(dtype)
    x = st.uniform(fs.undf).astype(dtype)
    ref = oracle.tri_solve(np.zeros(fs.undf, dtype), x, t0, tp, tm, fs.dofmap, nl)
    sp = api.Space(ctx, fs)
    y = torch.full((fs.undf,), float("nan"), dtype=torch.from_numpy(x).dtype, device=ctx.device)
    api.tri_solve(y, _t(x, ctx), _t(t0, ctx), _t(tp, ctx), _t(tm, ctx), sp)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)


def test_tri_solve_columns_not_at_a_constant_stride(ctx):
    """A W3 dofmap whose cells point at a permutation of the column blocks:
    the kernel reads each column's base from the map (a generated W3 map has
    its columns at a constant stride, and the base is formed instead)."""
    import dataclasses
    from lfric_core_amd import api
    print("case tri_solve permuted W3 columns C8L70 float32", flush=True)
    fs0 = build_dofmap(cubed_sphere(8), "W3", 70)
    perm = np.random.default_rng(5).permutation(fs0.dofmap.shape[0])
    fs = dataclasses.replace(fs0, dofmap=np.ascontiguousarray(fs0.dofmap[perm]))
    st = Stream(53)
    t0 = (4.0 + st.uniform(fs.undf)).astype(np.float32)
    tp = (0.5 * st.uniform(fs.undf)).astype(np.float32)
    tm = (0.5 * st.uniform(fs.undf)).astype(np.float32)
    x = st.uniform(fs.undf).astype(np.float32)
    ref = oracle.tri_solve(np.zeros(fs.undf, np.float32), x, t0, tp, tm, fs.dofmap, 70)
    sp = api.Space(ctx, fs)
    y = torch.full((fs.undf,), float("nan"), dtype=torch.float32, device=ctx.device)
    api.tri_solve(y, _t(x, ctx), _t(t0, ctx), _t(tp, ctx), _t(tm, ctx), sp)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("space,nc,nl", [("WTHETA", 3, 8), ("W2V", 6, 70), ("WTHETA", 5, 70)])
def test_operator_tri_solve_bitexact(ctx, dtype, space, nc, nl):
    from lfric_core_amd import api
    print(f"case operator_tri_solve {space} nc={nc} nl={nl} {dtype.__name__}", flush=True)
    fs = build_dofmap(cubed_sphere(nc), space, nl)
    assert fs.ndf == 2
    st = Stream(52)
    n3 = fs.mesh.ncells * nl
    m = np.empty((2, 2, n3))
    m[0, 0] = 2.0 + st.uniform(n3)
    m[0, 1] = 0.5 * st.uniform(n3)
    m[1, 0] = 0.5 * st.uniform(n3)
    m[1, 1] = 2.0 + st.uniform(n3)
    m = m.astype(dtype)
    x = st.uniform(fs.undf).astype(dtype)
    ref = oracle.operator_tri_solve(np.zeros(fs.undf, dtype), x, m, fs.dofmap, nl)
    sp = api.Space(ctx, fs)
    y = torch.full((fs.undf,), float("nan"), dtype=torch.from_numpy(x).dtype, device=ctx.device)
    api.operator_tri_solve(y, _t(x, ctx), _t(m, ctx), sp)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)


def test_tri_solve_bad_args(ctx):
    from lfric_core_amd import api
    from lfric_core_amd._ffi import LfrError
    print("case tri_solve errors", flush=True)
    fs = build_dofmap(cubed_sphere(2), "W2", 3)
    sp = api.Space(ctx, fs)
    z = torch.zeros(fs.undf, dtype=torch.float64, device=ctx.device)
    with pytest.raises(LfrError):   # not a W3 space
        api.tri_solve(z.clone(), z, z, z, z, sp)
    with pytest.raises(LfrError):   # not a 2-dof column space
        api.operator_tri_solve(z.clone(), z, torch.zeros((2, 2, fs.mesh.ncells * 3), dtype=torch.float64,
                                                         device=ctx.device), sp)


def test_operator_tri_solve_columns_not_at_a_constant_stride(ctx):
    """Wtheta with its cells pointing at a permutation of the column blocks:
    the column bases come from the map instead of the column stride."""
    import dataclasses
    from lfric_core_amd import api
    print("case operator_tri_solve permuted WTHETA columns C5L70 float32", flush=True)
    fs0 = build_dofmap(cubed_sphere(5), "WTHETA", 70)
    perm = np.random.default_rng(7).permutation(fs0.dofmap.shape[0])
    fs = dataclasses.replace(fs0, dofmap=np.ascontiguousarray(fs0.dofmap[perm]))
    st = Stream(54)
    n3 = fs.mesh.ncells * 70
    m = np.empty((2, 2, n3))
    m[0, 0] = 2.0 + st.uniform(n3)
    m[0, 1] = 0.5 * st.uniform(n3)
    m[1, 0] = 0.5 * st.uniform(n3)
    m[1, 1] = 2.0 + st.uniform(n3)
    m = m.astype(np.float32)
    x = st.uniform(fs.undf).astype(np.float32)
    ref = oracle.operator_tri_solve(np.zeros(fs.undf, np.float32), x, m, fs.dofmap, 70)
    sp = api.Space(ctx, fs)
    y = torch.full((fs.undf,), float("nan"), dtype=torch.float32, device=ctx.device)
    api.operator_tri_solve(y, _t(x, ctx), _t(m, ctx), sp)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
