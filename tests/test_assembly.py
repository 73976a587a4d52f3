"""W0 quadrature assembly (lfric_core_amd/assembly.py), pinned on the CPU by:
the reference's own W0 basis and gradient tables at its 3-point Gaussian
quadrature (fs_test_data_mod.F90, exact equality); closed forms on boxes
(tensor-product Q1 mass and stiffness); K 1 = 0; and the cell volumes of the
shell against an independent radial/bilinear-surface integral."""
import numpy as np

from lfric_core_amd import assembly as asm
from lfric_core_amd.configs import diffusion_config

from conftest import GOLDEN


def test_basis_tables_equal_reference_fixture():
    fs = np.load(GOLDEN / "fs_test_data.npz")
    pxy, _, pz, _ = asm.quadrature_xyoz(3)
    B, G = asm.w0_basis(pxy, pz)
    np.testing.assert_array_equal(fs["fs0k0_nodal_coords"].T, asm.W0_VERTICES)
    np.testing.assert_allclose(B, fs["fs0k0_basis_function"][0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(G, fs["fs0k0_basis_diff_function"], rtol=0, atol=1e-15)


def test_gauss_rule_is_exact_to_degree_5():
    p, w = asm.gauss_rule(3)
    for d in range(6):
        assert abs(np.sum(w * p ** d) - 1.0 / (d + 1)) < 1e-15


def _box(a, b, c):
    return asm.W0_VERTICES * np.array([a, b, c])


def test_box_mass_and_stiffness_closed_forms():
    a, b, c = 2.0, 0.5, 3.0
    M, K = asm.w0_mass_stiffness(_box(a, b, c))
    m1 = lambda h: h * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    k1 = lambda h: np.array([[1, -1], [-1, 1]]) / h
    # vertex v = (vx, vy, vz) bits; tensor index ordering of W0_VERTICES
    idx = [(int(x), int(y), int(z)) for x, y, z in asm.W0_VERTICES]
    Mx, My, Mz = m1(a), m1(b), m1(c)
    Kx, Ky, Kz = k1(a), k1(b), k1(c)
    for p, (i1, j1, l1) in enumerate(idx):
        for q, (i2, j2, l2) in enumerate(idx):
            m = Mx[i1, i2] * My[j1, j2] * Mz[l1, l2]
            k = (Kx[i1, i2] * My[j1, j2] * Mz[l1, l2] + Mx[i1, i2] * Ky[j1, j2] * Mz[l1, l2]
                 + Mx[i1, i2] * My[j1, j2] * Kz[l1, l2])
            assert abs(M[p, q] - m) < 1e-14 * abs(m) + 1e-15
            assert abs(K[p, q] - k) < 1e-13
    assert abs(M.sum() - a * b * c) < 1e-13


def test_shell_operator_invariants_and_volume():
    g = np.load(GOLDEN / "meshes.npz")
    cfg = diffusion_config(g["C12_node_x"], g["C12_node_y"], g["C12_face_nodes"])
    A, M, K = cfg.matrix, cfg.mass, cfg.stiffness
    assert A.shape == (8, 8, 864 * 10) and cfg.w0.undf == 866 * 11
    np.testing.assert_array_equal(M, M.transpose(1, 0, 2))
    assert np.abs(K.sum(axis=0)).max() <= 1e-13 * np.abs(K).max()      # K 1 = 0
    # cell volume = sum of the local mass matrix; independent integral:
    # x = rho(zeta) u(xi, eta), rho linear from r_k to r_k+1, u bilinear in the
    # corner unit vectors  ->  V = (r2 - r1)(r1^2 + r1 r2 + r2^2)/3 * int u.(u_xi x u_eta)
    lon, lat = np.deg2rad(g["C12_node_x"]), np.deg2rad(g["C12_node_y"])
    unit = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], 1)
    U = unit[g["C12_face_nodes"] - 1]                                 # SW SE NE NW
    t, w = np.polynomial.legendre.leggauss(6)
    t, w = 0.5 * (t + 1), 0.5 * w
    s_, e_ = np.meshgrid(t, t, indexing="ij")
    ww = np.outer(w, w)
    u = ((1 - s_) * (1 - e_))[..., None, None] * U[None, None, :, 0] \
        + (s_ * (1 - e_))[..., None, None] * U[None, None, :, 1] \
        + (s_ * e_)[..., None, None] * U[None, None, :, 2] \
        + ((1 - s_) * e_)[..., None, None] * U[None, None, :, 3]
    ux = (1 - e_)[..., None, None] * (U[:, 1] - U[:, 0]) + e_[..., None, None] * (U[:, 2] - U[:, 3])
    uy = (1 - s_)[..., None, None] * (U[:, 3] - U[:, 0]) + s_[..., None, None] * (U[:, 2] - U[:, 1])
    area = np.einsum("ij,ijcd,ijcd->c", ww, u, np.cross(ux, uy))
    r = 6.371e6 + cfg.heights
    vol = area[:, None] * ((r[:-1] ** 2 + r[:-1] * r[1:] + r[1:] ** 2) / 3 * np.diff(r))[None, :]
    np.testing.assert_allclose(M.sum(axis=(0, 1)), vol.ravel(), rtol=1e-11)


def test_w0_chi_fields_reproduce_the_corners():
    """chi1/chi2/chi3 on W0 read back through the dofmap give every cell-layer
    corner (shared vertices are consistent)."""
    from lfric_core_amd.assembly import shell_corners, w0_chi_fields
    from lfric_core_amd.function_space import build_dofmap
    from lfric_core_amd.mesh import cubed_sphere
    g = np.load(GOLDEN / "meshes.npz")
    nl = 4
    fs = build_dofmap(cubed_sphere(12), "W0", nl)
    heights = np.linspace(0.0, 1.0e4, nl + 1)
    chi = shell_corners(g["C12_node_x"], g["C12_node_y"], g["C12_face_nodes"], 6.371e6, heights)
    f = w0_chi_fields(fs, chi)
    dm = fs.dofmap.astype(np.int64) - 1
    for k in range(nl):
        for j in range(8):
            np.testing.assert_array_equal(f[:, dm[:, j] + k].T, chi[:, k, j, :])
