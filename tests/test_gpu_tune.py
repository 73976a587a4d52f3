"""The autotuner (lfr_tune_*): every launch-shape variant it can pick gives
the same bits, a first call stores its choice under the shape's key,
stored or forced choices are used, LFR_TUNE=0 times nothing, and the
choices survive a save/load round trip.  Outputs are checked against the
oracle for every forced variant."""
import numpy as np
import pytest

import oracle
from lfric_core_amd.function_space import build_dofmap
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _t(ctx, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)


def _appinv_case(ctx, nc=9):
    from lfric_core_amd import api
    m = cubed_sphere(nc)
    w3 = build_dofmap(m, "W3", 70)
    st = Stream(41)
    cm = np.empty((m.ncells, 70, 3))
    cm[:, :, 1] = 3.0 + st.uniform(m.ncells * 70).reshape(m.ncells, 70)
    cm[:, :, 0] = -1.0 + 0.1 * st.uniform(m.ncells * 70).reshape(m.ncells, 70)
    cm[:, :, 2] = -1.0 + 0.1 * st.uniform(m.ncells * 70).reshape(m.ncells, 70)
    cm[:, 0, 0] = 0.0
    cm[:, -1, 2] = 0.0
    cm = cm.astype(np.float32)
    x = st.uniform(w3.undf).astype(np.float32)
    ind = np.arange(1, 71, dtype=np.int32)
    ref = oracle.appinv(np.zeros_like(x), x, cm, w3.dofmap, ind)
    sp = api.Space(ctx, w3)
    op = api.CMAOperator(_t(ctx, cm), sp, ind)
    return m.ncells, op, _t(ctx, x), ref


def test_appinv_every_variant_identical(ctx):
    from lfric_core_amd import api
    print("case tune appinv C9L70", flush=True)
    ctx.tune_clear()
    ncell, op, x, ref = _appinv_case(ctx)
    key = f"appinv f32 nrow=70 ncell={ncell}"
    y = torch.empty_like(x)
    api.columnwise_op_appinv(y, x, op)          # first call: times the variants
    torch.cuda.synchronize()
    assert key in ctx.tune_entries()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
    for c in range(8):
        ctx.tune_set(key, c)
        y = torch.full_like(x, float("nan"))
        api.columnwise_op_appinv(y, x, op)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(y.cpu().numpy(), ref, err_msg=f"variant {c}")
        assert ctx.tune_get(key) == c


def test_tri_solve_every_variant_identical(ctx):
    from lfric_core_amd import api
    print("case tune tri_solve C9L70", flush=True)
    fs = build_dofmap(cubed_sphere(9), "W3", 70)
    st = Stream(43)
    u = fs.undf
    t0 = (4.0 + st.uniform(u)).astype(np.float32)
    tp = (-1.0 + 0.1 * st.uniform(u)).astype(np.float32)
    tm = (-1.0 + 0.1 * st.uniform(u)).astype(np.float32)
    x = st.uniform(u).astype(np.float32)
    ref = oracle.tri_solve(np.zeros(u, np.float32), x, t0, tp, tm, fs.dofmap, 70)
    sp = api.Space(ctx, fs)
    key = f"tri_solve f32 nl=70 ncell={fs.mesh.ncells}"
    ctx.tune_clear()
    args = [_t(ctx, a) for a in (x, t0, tp, tm)]
    for c in (None, 0, 1, 2, 3, 4, 5):
        if c is not None:
            ctx.tune_set(key, c)
        y = torch.full((u,), float("nan"), dtype=torch.float32, device=ctx.device)
        api.tri_solve(y, *args, sp)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(y.cpu().numpy(), ref, err_msg=f"variant {c}")
    assert ctx.tune_get(key) == 5


def test_operator_tri_solve_every_variant_identical(ctx):
    from lfric_core_amd import api
    print("case tune operator_tri_solve WTHETA C9L70", flush=True)
    fs = build_dofmap(cubed_sphere(9), "WTHETA", 70)
    st = Stream(47)
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
    key = f"operator_tri_solve f32 nl=70 ncell={fs.mesh.ncells}"
    ctx.tune_clear()
    for c in (None, 0, 1, 2):
        if c is not None:
            ctx.tune_set(key, c)
        y = torch.full((fs.undf,), float("nan"), dtype=torch.float32, device=ctx.device)
        api.operator_tri_solve(y, _t(ctx, x), _t(ctx, m), sp)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(y.cpu().numpy(), ref, err_msg=f"variant {c}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_packed_every_variant_identical(ctx, dtype):
    """The packed apply's variant is tuned when the operator is packed; each
    forced variant gives the oracle's bits."""
    from lfric_core_amd import api
    from lfric_core_amd.configs import precision_config
    print(f"case tune packed W2 C8L10 {dtype.__name__}", flush=True)
    cfg = precision_config(8, 10, with_div_grad=False)
    w2 = api.Space(ctx, cfg.w2)
    M = cfg.mass.astype(dtype)
    op = api.LMAOperator(_t(ctx, M), w2, w2, cfg.colouring)
    x = cfg.b.astype(dtype)
    ref = oracle.invoke_coloured(0, cfg.colouring, cfg.nlayers, np.zeros_like(x), x, M, cfg.w2.dofmap,
                                 cfg.w2.dofmap)
    ctx.tune_clear()
    pk = api.PackedOperator(op)
    keys = [k for k in ctx.tune_entries() if k.startswith("packed gather")]
    assert len(keys) == 1
    pk.close()
    for c in range(4):
        ctx.tune_set(keys[0], c)
        pk = api.PackedOperator(op)
        y = torch.full((cfg.w2.undf,), float("nan"), dtype=op.matrix.dtype, device=ctx.device)
        pk.apply(y, _t(ctx, x), overwrite=True)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(y.cpu().numpy(), ref, err_msg=f"variant {c}")
        pk.close()


@pytest.mark.parametrize("mprec", [4, 2])
def test_multi16_owner_every_variant_identical(ctx, mprec):
    """The 16-field owner-computes apply: its prefetch-depth variants give the
    same bits; the first overwrite apply stores the choice, an increment
    apply uses it without timing."""
    from lfric_core_amd import api
    from lfric_core_amd import function_space as fsm
    from lfric_core_amd.configs import spd_lma
    print(f"case tune multi16 owner W2k1 C4L20 mprec={mprec}", flush=True)
    mesh = cubed_sphere(4)
    nl = 20
    fs = fsm.build_dofmap(mesh, "W2", nl, k=1)
    col = fsm.colour_cells(mesh)
    st = Stream(53)
    M = spd_lma(st, mesh.ncells * nl, 36, 36.0, std=1.0 / 6.0).astype(np.float32)
    X = st.uniform(fs.undf * 16).reshape(fs.undf, 16).astype(np.float32)
    sp = api.Space(ctx, fs)
    op = api.LMAOperator(_t(ctx, M), sp, sp, col)
    mop = api.MultiOperator(op, mprec)
    assert mop.info()["launches"] == 1
    key = f"multi16 owner prec={mprec} nl={nl} ncell={mesh.ncells}"
    ctx.tune_clear()
    x = _t(ctx, X)
    y0 = torch.full((fs.undf, 16), float("nan"), dtype=torch.float32, device=ctx.device)
    mop.apply(y0, x, overwrite=True)
    torch.cuda.synchronize()
    assert key in ctx.tune_entries()
    for c in (0, 1):
        ctx.tune_set(key, c)
        y = torch.full_like(y0, float("nan"))
        mop.apply(y, x, overwrite=True)
        assert torch.equal(y, y0), f"variant {c}"
        y.copy_(y0)
        mop.apply(y, x, overwrite=False)
        torch.testing.assert_close(y, 2 * y0, rtol=1e-6, atol=1e-6 * float(y0.abs().max()))
    ctx.tune_clear()
    y = y0.clone()
    mop.apply(y, x, overwrite=False)           # increment form: no timing
    assert ctx.tune_entries() == {}
    mop.close()


def test_multigrid_levels_tuned_and_no_tuning_switch(ctx, monkeypatch):
    from lfric_core_amd.multigrid import Multigrid, build_hierarchy
    print("case tune multigrid keys, LFR_TUNE=0", flush=True)
    H = build_hierarchy((8, 4), nlayers=10)
    ctx.tune_clear()
    mg = Multigrid(ctx, H, prec=4, nrhs=2)
    keys = [k for k in ctx.tune_entries() if k.startswith("mg smooth")]
    assert len(keys) == 2
    assert all(b in (64, 128, 256) for b in mg.block_sizes)
    mg.close()
    ctx.tune_clear()
    monkeypatch.setenv("LFR_TUNE", "0")
    mg = Multigrid(ctx, H, prec=4, nrhs=2)
    assert mg.block_sizes == [256, 256] and ctx.tune_entries() == {}
    mg.close()
    ncell, op, x, ref = _appinv_case(ctx, nc=6)
    from lfric_core_amd import api
    y = torch.empty_like(x)
    api.columnwise_op_appinv(y, x, op)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
    assert ctx.tune_entries() == {}


def test_save_load_round_trip(ctx, tmp_path):
    print("case tune save/load", flush=True)
    ctx.tune_clear()
    ctx.tune_set("appinv f32 nrow=70 ncell=486", 1)
    ctx.tune_set("mg smooth prec=4 nl=10 ncell=384 rhs=2", 0)
    before = ctx.tune_entries()
    p = tmp_path / "tune.txt"
    ctx.tune_save(p)
    ctx.tune_clear()
    assert ctx.tune_entries() == {}
    ctx.tune_load(p)
    assert ctx.tune_entries() == before
    from lfric_core_amd._ffi import LfrError
    with pytest.raises(LfrError):
        ctx.tune_get("no such key")
    ctx.tune_clear()
