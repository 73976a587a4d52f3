"""lfr_restrict_scalar / lfr_prolong_scalar on the device against the oracle
(itself bit-exact against the flang-built reference kernels), bit for bit,
with the reference's intergrid maps; plus the argument checks."""

import numpy as np
import pytest

import oracle
from lfric_core_amd.function_space import build_dofmap
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("space", ["W3", "WTHETA"])
@pytest.mark.parametrize("key,nc,nf,nl", [("C8_C16_map", 8, 16, 70), ("C4_C12_map", 4, 12, 5)])
def test_transfers_bitexact(ctx, key, nc, nf, nl, space, dtype):
    from lfric_core_amd import api
    print(f"case intergrid {key} {space} nl={nl} {dtype.__name__}", flush=True)
    cm = np.load(GOLDEN / "meshes.npz")[key]
    fc = build_dofmap(cubed_sphere(nc), space, nl)
    ff = build_dofmap(cubed_sphere(nf), space, nl)
    sc, sf = api.Space(ctx, fc), api.Space(ctx, ff)
    st = Stream(63)
    fine = st.uniform(ff.undf).astype(dtype)
    coarse = st.uniform(fc.undf).astype(dtype)
    ref_r = oracle.restrict_scalar(np.full(fc.undf, np.nan, dtype), fine, fc.dofmap, ff.dofmap, cm, nl)
    ref_p = oracle.prolong_scalar(fine.copy(), coarse, fc.dofmap, ff.dofmap, cm, nl)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)
    ig = api.Intergrid(sc, sf, cm)
    yc = torch.full((fc.undf,), float("nan"), dtype=t(fine).dtype, device=ctx.device)
    api.restrict_scalar(yc, t(fine), ig)
    yf = t(fine.copy())
    api.prolong_scalar(yf, t(coarse), ig)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(yc.cpu().numpy(), ref_r)
    np.testing.assert_array_equal(yf.cpu().numpy(), ref_p)


def test_transfer_bad_args(ctx):
    from lfric_core_amd import api
    from lfric_core_amd._ffi import LfrError
    print("case intergrid errors", flush=True)
    cm = np.load(GOLDEN / "meshes.npz")["C8_C16_map"]
    fc = build_dofmap(cubed_sphere(8), "W3", 3)
    ff = build_dofmap(cubed_sphere(16), "W3", 3)
    sc, sf = api.Space(ctx, fc), api.Space(ctx, ff)
    z = lambda n: torch.zeros(n, dtype=torch.float64, device=ctx.device)
    bad = cm.copy()
    bad[0, 0, 0] = ff.mesh.ncells + 1
    with pytest.raises(LfrError):   # a fine cell out of range
        api.Intergrid(sc, sf, bad)
    with pytest.raises(LfrError):   # the map does not cover the fine mesh
        api.Intergrid(sc, sf, cm[:, :1, :])
    w2 = build_dofmap(cubed_sphere(8), "W2", 3)
    with pytest.raises(LfrError):   # not a scalar space
        api.Intergrid(api.Space(ctx, w2), sf, cm)
    ig = api.Intergrid(sc, sf, cm)
    with pytest.raises(LfrError):   # field on the wrong space
        api.restrict_scalar(z(ff.undf), z(ff.undf), ig)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("key,nc,nf,nl", [("C8_C16_map", 8, 16, 70), ("C4_C12_map", 4, 12, 5)])
def test_restrict_w2_bitexact(ctx, key, nc, nf, nl, dtype):
    from lfric_core_amd import api
    print(f"case restrict_w2 {key} nl={nl} {dtype.__name__}", flush=True)
    cm = np.load(GOLDEN / "meshes.npz")[key]
    fc = build_dofmap(cubed_sphere(nc), "W2", nl)
    ff = build_dofmap(cubed_sphere(nf), "W2", nl)
    fine = Stream(65).uniform(ff.undf).astype(dtype)
    ref = oracle.restrict_w2(np.full(fc.undf, np.nan, dtype), fine, fc.dofmap, ff.dofmap, cm, nl, fc.undf)
    ig = api.Intergrid(api.Space(ctx, fc), api.Space(ctx, ff), cm)
    y = torch.full((fc.undf,), float("nan"), dtype=torch.from_numpy(fine).dtype, device=ctx.device)
    api.restrict_w2(y, torch.from_numpy(fine).to(ctx.device), ig)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
    from lfric_core_amd._ffi import LfrError
    with pytest.raises(LfrError):   # the scalar transfer on a W2 map
        api.restrict_scalar(y, torch.from_numpy(fine).to(ctx.device), ig)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("key,nc,nf,nl", [("C8_C16_map", 8, 16, 70), ("C4_C12_map", 4, 12, 5)])
def test_prolong_w2_bitexact(ctx, key, nc, nf, nl, dtype):
    from lfric_core_amd import api
    print(f"case prolong_w2 {key} nl={nl} {dtype.__name__}", flush=True)
    cm = np.load(GOLDEN / "meshes.npz")[key]
    fc = build_dofmap(cubed_sphere(nc), "W2", nl)
    ff = build_dofmap(cubed_sphere(nf), "W2", nl)
    u = Stream(68).uniform(fc.undf).astype(dtype)
    ref = oracle.prolong_w2(np.full(ff.undf, np.nan, dtype), u, fc.dofmap, ff.dofmap, cm, nl)
    ig = api.Intergrid(api.Space(ctx, fc), api.Space(ctx, ff), cm)
    y = torch.full((ff.undf,), float("nan"), dtype=torch.from_numpy(u).dtype, device=ctx.device)
    api.prolong_w2(y, torch.from_numpy(u).to(ctx.device), ig)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
