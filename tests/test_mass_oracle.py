"""The mass-matrix oracle (oracle/assembly.py) pinned by closed forms and by
the independent host W0 assembly, with the reference's own basis tables
(infrastructure/unit-test/support/fs_test_data_mod.F90: lowest-order spaces
at the 3x3x3 Gauss points, horizontal points x outer, y inner)."""

import numpy as np
import pytest

from oracle.assembly import mass_matrix
from lfric_core_amd.assembly import W0_VERTICES, gauss_rule, shell_corners, w0_mass_stiffness

from conftest import GOLDEN


@pytest.fixture(scope="module")
def tables():
    d = np.load(GOLDEN / "fs_test_data.npz")
    _, w = gauss_rule(3)
    wh = np.outer(w, w).ravel()
    return d, wh, w


def _affine(A, b, n=1):
    """chi_e (n, 3, 8) of the cell xi -> A xi + b."""
    c = (W0_VERTICES @ np.asarray(A, np.float64).T) + np.asarray(b, np.float64)    # (8, 3)
    return np.broadcast_to(c.T, (n, 3, 8)).copy()


def test_golden_tables_are_the_gauss_point_w0_basis(tables):
    from lfric_core_amd.assembly import quadrature_xyoz, w0_basis
    d, _, _ = tables
    pxy, _, pz, _ = quadrature_xyoz(3)
    B, G = w0_basis(pxy, pz)
    np.testing.assert_array_equal(d["fs0k0_basis_function"][0], B)
    np.testing.assert_array_equal(d["fs0k0_basis_diff_function"], G)


def test_w0_mass_matches_host_assembly_on_the_shell(tables):
    d, wh, wv = tables
    g = np.load(GOLDEN / "meshes.npz")
    heights = np.linspace(0.0, 3.0e4, 4)
    chi = shell_corners(g["C12_node_x"], g["C12_node_y"], g["C12_face_nodes"], 6.371e6, heights)[:40]
    chi_e = chi.reshape(-1, 8, 3).transpose(0, 2, 1)
    mm = mass_matrix(0, d["fs0k0_basis_function"], d["fs0k0_basis_diff_function"], wh, wv, chi_e)
    host, _ = w0_mass_stiffness(chi, 3)
    host = host.reshape(-1, 8, 8)
    assert np.abs(mm - host).max() <= 1e-13 * np.abs(host).max()


def test_w3_mass_is_the_cell_volume(tables):
    d, wh, wv = tables
    A = np.array([[2.0, 0.3, 0.1], [-0.2, 1.5, 0.4], [0.05, 0.1, 0.7]])
    mm = mass_matrix(0, d["fs3k0_basis_function"], d["fs0k0_basis_diff_function"], wh, wv,
                     _affine(A, [1.0, -2.0, 3.0]))
    assert mm.shape == (1, 1, 1)
    assert mm[0, 0, 0] == pytest.approx(np.linalg.det(A), rel=1e-14)


def test_wtheta_mass_on_a_box(tables):
    d, wh, wv = tables
    a, b, c = 3.0, 0.5, 2.0
    mm = mass_matrix(0, d["fstk0_basis_function"], d["fs0k0_basis_diff_function"], wh, wv,
                     _affine(np.diag([a, b, c]), [0.0, 0.0, 0.0]))
    # basis 1 - z (bottom) and z (top): integrals 1/3 and 1/6
    ref = a * b * c * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    np.testing.assert_allclose(mm[0], ref, rtol=1e-14)


def test_w2_mass_on_a_box(tables):
    d, wh, wv = tables
    a, b, c = 3.0, 0.5, 2.0
    mm = mass_matrix(1, d["fs2k0_basis_function"], d["fs0k0_basis_diff_function"], wh, wv,
                     _affine(np.diag([a, b, c]), [5.0, 1.0, -1.0]))[0]
    # dofs W, S, E, N, B, T: normal components (1-x, 0, 0), (0, -(1-y), 0),
    # (x, 0, 0), (0, -y, 0), (0, 0, 1-z), (0, 0, z); Piola: |J e_d|^2 / det J
    ref = np.zeros((6, 6))
    for (p, q), s in zip(((0, 2), (1, 3), (4, 5)), (a / (b * c), b / (a * c), c / (a * b))):
        ref[p, p] = ref[q, q] = s / 3
        ref[p, q] = ref[q, p] = s / 6
    np.testing.assert_allclose(mm, ref, rtol=1e-14, atol=1e-15)


def test_w2_mass_on_sheared_cells(tables):
    """Affine cells: M_ab = (1/det A) sum_ij (A^T A)_ij int v_a,i v_b,j, the
    integrals from the tables alone (a transposed Jacobian or a product in
    place of the quotient fails this for a non-symmetric A)."""
    d, wh, wv = tables
    basis = d["fs2k0_basis_function"]
    w = wh[:, None] * wv[None, :]
    I = np.einsum("iahv,jbhv,hv->abij", basis, basis, w)
    for A in (np.array([[2.0, 0.6, 0.1], [-0.3, 1.5, 0.4], [0.2, -0.1, 0.7]]),
              np.array([[1.0, 0.0, 0.9], [0.0, 1.0, 0.0], [0.0, 0.5, 1.0]])):
        mm = mass_matrix(1, basis, d["fs0k0_basis_diff_function"], wh, wv, _affine(A, [0.5, 0.5, 0.5]))[0]
        ref = np.einsum("ij,abij->ab", A.T @ A, I) / np.linalg.det(A)
        np.testing.assert_allclose(mm, ref, rtol=1e-13, atol=1e-15)


def test_w2h_w2v_masses_are_the_w2_blocks(tables):
    d, wh, wv = tables
    A = np.array([[2.0, 0.6, 0.1], [-0.3, 1.5, 0.4], [0.2, -0.1, 0.7]])
    chi_e = _affine(A, [0.0, 0.0, 0.0])
    dchi = d["fs0k0_basis_diff_function"]
    full = mass_matrix(1, d["fs2k0_basis_function"], dchi, wh, wv, chi_e)[0]
    h = mass_matrix(1, d["fs2hk0_basis_function"], dchi, wh, wv, chi_e)[0]
    v = mass_matrix(1, d["fs2vk0_basis_function"], dchi, wh, wv, chi_e)[0]
    np.testing.assert_array_equal(h, full[:4, :4])
    np.testing.assert_array_equal(v, full[4:, 4:])


def _derham_tables(d, wh, wv):
    return {"wqp_h": wh, "wqp_v": wv, "diff_basis_chi": d["fs0k0_basis_diff_function"],
            "basis_w0": d["fs0k0_basis_function"], "diff_basis_w0": d["fs0k0_basis_diff_function"],
            "basis_w1": d["fs1k0_basis_function"], "diff_basis_w1": d["fs1k0_basis_diff_function"],
            "basis_w2": d["fs2k0_basis_function"], "diff_basis_w2": d["fs2k0_basis_diff_function"],
            # W2broken has the W2 local basis (discontinuous only across cells)
            "basis_w2b": d["fs2k0_basis_function"], "diff_basis_w2b": d["fs2k0_basis_diff_function"],
            "basis_w3": d["fs3k0_basis_function"], "basis_wt": d["fstk0_basis_function"]}


def _represent(target, basis):
    """Integer coefficients K with target(:, b, q) = sum_e K[e, b] basis(:, e, q)
    at every point (least squares over the points, then rounded and checked)."""
    nb = target.shape[1]
    A = np.moveaxis(basis, 1, -1).reshape(-1, basis.shape[1])
    B = np.moveaxis(target, 1, -1).reshape(-1, nb)
    K, *_ = np.linalg.lstsq(A, B, rcond=None)
    Ki = np.rint(K)
    assert np.abs(A @ Ki - B).max() < 1e-14
    return Ki


def _incidence(d):
    G = _represent(d["fs0k0_basis_diff_function"], d["fs1k0_basis_function"])     # (12, 8)
    C = _represent(d["fs1k0_basis_diff_function"], d["fs2k0_basis_function"])     # (6, 12)
    D = _represent(d["fs2k0_basis_diff_function"], d["fs3k0_basis_function"])     # (1, 6)
    return G, C, D


def test_the_tables_form_an_exact_sequence(tables):
    d, _, _ = tables
    G, C, D = _incidence(d)
    assert set(np.unique(G)) <= {-1.0, 0.0, 1.0} and np.abs(G).sum(axis=0).tolist() == [3.0] * 8
    assert np.abs(G).sum(axis=1).tolist() == [2.0] * 12       # an edge has two ends
    assert np.abs(C).sum(axis=1).tolist() == [4.0] * 6        # a face has four edges
    assert np.abs(D).tolist() == [[1.0] * 6]
    assert not (C @ G).any() and not (D @ C).any()


def test_derham_masses_equal_the_mass_kernels(tables):
    from oracle.assembly import derham_matrices
    d, wh, wv = tables
    g = np.load(GOLDEN / "meshes.npz")
    chi = shell_corners(g["C12_node_x"], g["C12_node_y"], g["C12_face_nodes"], 6.371e6,
                        np.linspace(0.0, 3.0e4, 3))[:20]
    chi_e = chi.reshape(-1, 8, 3).transpose(0, 2, 1)
    out = derham_matrices(_derham_tables(d, wh, wv), chi_e)
    dchi = d["fs0k0_basis_diff_function"]
    for key, name, kind in (("mm0", "fs0k0", 0), ("mm2", "fs2k0", 1), ("mm2b", "fs2k0", 1),
                            ("mm3", "fs3k0", 0), ("mmt", "fstk0", 0)):
        np.testing.assert_array_equal(out[key], mass_matrix(kind, d[f"{name}_basis_function"], dchi, wh, wv,
                                                            chi_e), err_msg=key)


def test_w1_mass_on_sheared_cells(tables):
    from oracle.assembly import form
    d, wh, wv = tables
    basis = d["fs1k0_basis_function"]
    w = wh[:, None] * wv[None, :]
    I = np.einsum("iahv,jbhv,hv->abij", basis, basis, w)
    A = np.array([[2.0, 0.6, 0.1], [-0.3, 1.5, 0.4], [0.2, -0.1, 0.7]])
    mm = form("cov", basis, basis, d["fs0k0_basis_diff_function"], wh, wv, _affine(A, [1.0, 2.0, 3.0]),
              symmetric=True)[0]
    Ai = np.linalg.inv(A)
    ref = np.linalg.det(A) * np.einsum("ij,abij->ab", Ai @ Ai.T, I)
    np.testing.assert_allclose(mm, ref, rtol=1e-13, atol=1e-15)


def test_grad_curl_div_through_the_incidence_matrices(tables):
    """On curved shell cells: grad = M1 G, curl = M2 C, div = D (the
    derivative of a lowest-order basis function lies in the next space)."""
    from oracle.assembly import derham_matrices
    d, wh, wv = tables
    G, C, D = _incidence(d)
    g = np.load(GOLDEN / "meshes.npz")
    chi = shell_corners(g["C12_node_x"], g["C12_node_y"], g["C12_face_nodes"], 6.371e6,
                        np.array([0.0, 1.0e3, 5.0e3]))[100:130]
    chi_e = chi.reshape(-1, 8, 3).transpose(0, 2, 1)
    out = derham_matrices(_derham_tables(d, wh, wv), chi_e)
    rg = out["mm1"] @ G
    np.testing.assert_allclose(out["grad"], rg, rtol=0, atol=1e-12 * np.abs(rg).max())
    rc = out["mm2"] @ C
    np.testing.assert_allclose(out["curl"], rc, rtol=0, atol=1e-12 * np.abs(rc).max())
    np.testing.assert_allclose(out["div"], np.broadcast_to(D, out["div"].shape), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(out["broken_div"], out["div"])
