"""Device V-cycle (lfr_mg_*) against the numpy restatement in oracle/multigrid.py.

Every kernel of the cycle is elementwise in a fixed operation order (no
reductions), so the iterate after each cycle is bit-exact against the oracle
in FP64 and in FP32.  Residual norms are reductions in a different order:
relative tolerance 1e-12 (FP64) / 1e-5 (FP32).
"""
import numpy as np
import pytest

from lfric_core_amd.multigrid import Multigrid, build_hierarchy
from lfric_core_amd.rng import Stream

pytestmark = pytest.mark.gpu


def _run(ctx, H, prec, nrhs, ncycles, seed=11, **kw):
    import torch
    from oracle import multigrid as omg
    dt = torch.float32 if prec == 4 else torch.float64
    npdt = np.float32 if prec == 4 else np.float64
    st = Stream(seed)
    b = st.uniform(nrhs * H.fine_undf).reshape(nrhs, -1).astype(npdt)
    x0 = (0.1 * st.uniform(nrhs * H.fine_undf)).reshape(nrhs, -1).astype(npdt)
    mg = Multigrid(ctx, H, prec=prec, nrhs=nrhs, **kw)
    xd = torch.from_numpy(x0.copy()).to(ctx.device)
    bd = torch.from_numpy(b).to(ctx.device)
    res = mg.solve(xd, bd, ncycles)
    torch.cuda.synchronize()
    okw = {k: kw[k] for k in ("omega", "nu1", "nu2", "ncoarse") if k in kw}
    xo, reso = omg.solve(H, x0, b, ncycles, dtype=npdt, **okw)
    mg.close()
    return xd.cpu().numpy(), res, xo, reso


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nl", [10, 6])
def test_vcycle_bit_exact(ctx, prec, nl):
    H = build_hierarchy((16, 8, 4), nlayers=nl)
    print(f"mg prec={prec} nl={nl}", flush=True)
    x, res, xo, reso = _run(ctx, H, prec, nrhs=4, ncycles=3)
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_allclose(res, reso, rtol=1e-12 if prec == 8 else 1e-5)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("nrhs", [8, 32])
def test_vcycle_70_layers_bit_exact(ctx, prec, nrhs):
    """70 layers: the unrolled smoother (backward-sweep loads a chunk ahead)
    and, for up to 16 right-hand sides, the residual-and-restriction with one
    lane per fine child (children summed on the child-0 lane); for 32 the
    loop form.  Every right-hand side bit-exact against the oracle."""
    H = build_hierarchy((8, 4), nlayers=70)
    print(f"mg L70 prec={prec} nrhs={nrhs}", flush=True)
    x, res, xo, reso = _run(ctx, H, prec, nrhs=nrhs, ncycles=2)
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_allclose(res, reso, rtol=1e-12 if prec == 8 else 1e-5)


def test_vcycle_options_and_repeat(ctx):
    """Other smoothing counts; a second solve replays the captured cycle."""
    H = build_hierarchy((8, 4, 2), nlayers=10)
    kw = dict(omega=0.7, nu1=4, nu2=2, ncoarse=4)
    x, res, xo, reso = _run(ctx, H, 8, nrhs=2, ncycles=2, **kw)
    np.testing.assert_array_equal(x, xo)
    x2, _, xo2, _ = _run(ctx, H, 8, nrhs=2, ncycles=4, seed=12, **kw)
    np.testing.assert_array_equal(x2, xo2)


def test_c48_convergence_fp32_vs_fp64(ctx):
    """Full-depth hierarchy below the workload: FP32 and FP64 residual histories agree
    while above FP32 round-off, and each cycle reduces the residual."""
    H = build_hierarchy((48, 24, 12, 6), nlayers=70)
    import torch
    st = Stream(5)
    b = st.uniform(2 * H.fine_undf).reshape(2, -1)
    out = {}
    for prec, dt in ((8, torch.float64), (4, torch.float32)):
        mg = Multigrid(ctx, H, prec=prec, nrhs=2)
        xd = torch.zeros(2, H.fine_undf, dtype=dt, device=ctx.device)
        out[prec] = mg.solve(xd, torch.from_numpy(b).to(ctx.device, dt), 6)
        mg.close()
    r64, r32 = out[8], out[4]
    rate = (r64[-1] / r64[0]) ** (1 / 6)
    assert np.all(rate < 0.6), rate
    assert np.all(np.diff(r64, axis=0) < 0)
    np.testing.assert_allclose(r32[:4], r64[:4], rtol=1e-3)


def test_bad_arguments(ctx):
    import torch
    from lfric_core_amd._ffi import LfrError
    H = build_hierarchy((8, 4), nlayers=10)
    mg = Multigrid(ctx, H, prec=8, nrhs=1)
    with pytest.raises(LfrError):
        mg.solve(torch.zeros(2, H.fine_undf, dtype=torch.float64, device=ctx.device),
                 torch.zeros(2, H.fine_undf, dtype=torch.float64, device=ctx.device))
    mg.close()
    with pytest.raises(LfrError):
        Multigrid(ctx, H, prec=8, nrhs=3)              # right-hand sides: a power of 2
    H.levels[1].children[0, 0] = H.levels[1].children[0, 1]
    with pytest.raises(LfrError):
        Multigrid(ctx, H, prec=8, nrhs=1)
