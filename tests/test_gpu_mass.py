"""lfr_compute_mass_matrix and lfr_compute_derham_matrices on the device
against the assembly oracle (oracle/assembly.py), on
the C12 shell with 5 layers, with the reference's own basis tables
(infrastructure/unit-test/support/fs_test_data_mod.F90) for W0, Wtheta, W3
(scalar kernel) and W2, W2H, W2V (Piola kernel).  Same per-entry summation
order as the oracle; the Jacobian products may round differently (no FMA on
either side, but numpy evaluates its broadcasts in its own order), so the
bound is 1e-13 of the largest entry, and the result is exactly symmetric."""

import numpy as np
import pytest

from oracle.assembly import mass_matrix
from lfric_core_amd.assembly import gauss_rule, shell_corners, w0_chi_fields
from lfric_core_amd.function_space import build_dofmap
from lfric_core_amd.mesh import cubed_sphere

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

NL = 5


@pytest.fixture(scope="module")
def shell():
    g = np.load(GOLDEN / "meshes.npz")
    heights = np.linspace(0.0, 2.0e4, NL + 1)
    chi = shell_corners(g["C12_node_x"], g["C12_node_y"], g["C12_face_nodes"], 6.371e6, heights)
    fs = build_dofmap(cubed_sphere(12), "W0", NL)
    return chi, fs


SPACES = [("fs0k0", 0), ("fstk0", 0), ("fs3k0", 0), ("fs2k0", 1), ("fs2hk0", 1), ("fs2vk0", 1)]


@pytest.mark.parametrize("name,kind", SPACES)
def test_device_mass_matches_oracle(ctx, shell, name, kind):
    from lfric_core_amd import api
    print(f"case device mass {name} kind={kind} C12L{NL}", flush=True)
    chi, fs = shell
    d = np.load(GOLDEN / "fs_test_data.npz")
    _, w = gauss_rule(3)
    wh = np.outer(w, w).ravel()
    basis = d[f"{name}_basis_function"]
    dchi = d["fs0k0_basis_diff_function"]
    chi_e = chi.reshape(-1, 8, 3).transpose(0, 2, 1)          # (ncell_3d, 3, 8), ik = cell*nl + k
    ref = mass_matrix(kind, basis, dchi, wh, w, chi_e)        # (ik, df, df2)
    ref_ls = np.ascontiguousarray(ref.transpose(2, 1, 0))     # local_stencil (ndf, ndf, n3) order
    sp = api.Space(ctx, fs)
    chi_f = torch.from_numpy(w0_chi_fields(fs, chi)).to(ctx.device)
    mm = api.compute_mass_matrix(kind, basis, dchi, wh, w, chi_f, sp)
    torch.cuda.synchronize()
    got = mm.cpu().numpy()
    assert got.shape == ref_ls.shape
    assert np.abs(got - ref_ls).max() <= 1e-13 * np.abs(ref_ls).max()
    np.testing.assert_array_equal(got, got.transpose(1, 0, 2))


def test_device_mass_w0_equals_w0_entry_point(ctx, shell):
    """The table-driven scalar kernel with the W0 tables and the built-in W0
    assembler (its own Gauss rule and trilinear basis) agree to rounding."""
    from lfric_core_amd import api
    print("case device mass W0 tables vs lfr_compute_mass_matrix_w0", flush=True)
    chi, fs = shell
    d = np.load(GOLDEN / "fs_test_data.npz")
    _, w = gauss_rule(3)
    wh = np.outer(w, w).ravel()
    sp = api.Space(ctx, fs)
    chi_f = torch.from_numpy(w0_chi_fields(fs, chi)).to(ctx.device)
    a = api.compute_mass_matrix(0, d["fs0k0_basis_function"], d["fs0k0_basis_diff_function"], wh, w, chi_f, sp)
    b = api.compute_mass_matrix_w0(chi_f, sp, nq=3)
    torch.cuda.synchronize()
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_device_mass_bad_args(ctx, shell):
    from lfric_core_amd import api
    from lfric_core_amd._ffi import LfrError
    print("case device mass errors", flush=True)
    chi, fs = shell
    d = np.load(GOLDEN / "fs_test_data.npz")
    _, w = gauss_rule(3)
    wh = np.outer(w, w).ravel()
    sp = api.Space(ctx, fs)
    chi_f = torch.from_numpy(w0_chi_fields(fs, chi)).to(ctx.device)
    dchi = d["fs0k0_basis_diff_function"]
    with pytest.raises(LfrError):   # W2 tables with the scalar kernel
        api.compute_mass_matrix(0, d["fs2k0_basis_function"], dchi, wh, w, chi_f, sp)
    with pytest.raises(LfrError):   # no Piola kernel for 12 dofs
        api.compute_mass_matrix(1, d["fs1k0_basis_function"], dchi, wh, w, chi_f, sp)
    with pytest.raises(LfrError):   # chi not on a W0 space
        w3 = api.Space(ctx, build_dofmap(cubed_sphere(12), "W3", NL))
        api.compute_mass_matrix(0, d["fs3k0_basis_function"], dchi, wh, w,
                                torch.zeros((3, w3.undf), dtype=torch.float64, device=ctx.device), w3)


def _derham_tables(d, wh, wv):
    return {"wqp_h": wh, "wqp_v": wv, "diff_basis_chi": d["fs0k0_basis_diff_function"],
            "basis_w0": d["fs0k0_basis_function"], "diff_basis_w0": d["fs0k0_basis_diff_function"],
            "basis_w1": d["fs1k0_basis_function"], "diff_basis_w1": d["fs1k0_basis_diff_function"],
            "basis_w2": d["fs2k0_basis_function"], "diff_basis_w2": d["fs2k0_basis_diff_function"],
            "basis_w2b": d["fs2k0_basis_function"], "diff_basis_w2b": d["fs2k0_basis_diff_function"],
            "basis_w3": d["fs3k0_basis_function"], "basis_wt": d["fstk0_basis_function"]}


def test_device_derham_matrices_match_oracle(ctx, shell):
    """All ten outputs of compute_derham_matrices on the C12 shell against the
    oracle; the mass outputs also equal lfr_compute_mass_matrix bit for bit
    (one kernel template)."""
    from lfric_core_amd import api
    from oracle.assembly import derham_matrices
    print(f"case device de Rham matrices C12L{NL}", flush=True)
    chi, fs = shell
    d = np.load(GOLDEN / "fs_test_data.npz")
    _, w = gauss_rule(3)
    wh = np.outer(w, w).ravel()
    t = _derham_tables(d, wh, w)
    chi_e = chi.reshape(-1, 8, 3).transpose(0, 2, 1)
    ref = derham_matrices(t, chi_e)
    sp = api.Space(ctx, fs)
    chi_f = torch.from_numpy(w0_chi_fields(fs, chi)).to(ctx.device)
    out = api.compute_derham_matrices(t, chi_f, sp)
    m2 = api.compute_mass_matrix(1, d["fs2k0_basis_function"], d["fs0k0_basis_diff_function"], wh, w, chi_f, sp)
    torch.cuda.synchronize()
    for name, r in ref.items():
        got = out[name].cpu().numpy()
        r_ls = np.ascontiguousarray(r.transpose(2, 1, 0))
        assert got.shape == r_ls.shape, name
        err = np.abs(got - r_ls).max() / np.abs(r_ls).max()
        print(f"  {name}: max error / max entry {err:.2e}", flush=True)
        assert err <= 1e-13, name
    assert torch.equal(out["mm2"], m2)
    assert torch.equal(out["mm2b"], m2)
    # a subset of outputs, and the error when a needed table is missing
    sub = api.compute_derham_matrices({k: v for k, v in t.items() if k not in ("basis_w1", "diff_basis_w1")},
                                      chi_f, sp, outputs=("div", "mm3"))
    torch.cuda.synchronize()
    assert set(sub) == {"div", "mm3"} and torch.equal(sub["div"], out["div"])
    from lfric_core_amd._ffi import LfrError
    with pytest.raises((LfrError, KeyError)):
        api.compute_derham_matrices({k: v for k, v in t.items() if k != "basis_w1"}, chi_f, sp,
                                    outputs=("grad",))
