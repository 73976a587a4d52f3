"""Host (FP64) assembly of lowest-order W0 local matrices by Gaussian quadrature.

Used to build the assembled-operator configuration (C12 shell, W0 mass +
1e-3 stiffness).  Restated pieces:

* Gauss-Legendre rule on [0, 1]
  (infrastructure/source/quadrature/quadrature_rule_gaussian_mod.F90, the
  published Legendre-root rule; roots taken here from numpy's ``leggauss``);
* horizontal points in the order of quadrature_xyoz_mod.F90:235-242
  (x outer, y inner, weight w_x * w_y);
* W0 basis at the reference-cube vertices in the order of the reference
  element (bottom SW, SE, NE, NW, then top), trilinear;
* coordinate Jacobian with Cartesian (X, Y, Z) coordinates
  (components/science/source/kernel/geometry/sci_coordinate_jacobian_mod.F90:330-345,
  the ``coord_system_xyz`` branch, and the cofactor determinant :430-440);
* mass matrix of sci_compute_mass_matrix_kernel_w_scalar_mod.F90:150-174
  (quadrature sum qp_v outer, qp_h inner, terms w_h w_v phi_i phi_j detJ,
  lower triangle computed and mirrored).

The stiffness matrix has no kernel in the reference: it is defined here as
K_ij = sum_q w_q (J^-T grad phi_i) . (J^-T grad phi_j) |det J|.

Geometry: the cells' corner longitudes/latitudes come from the cubed-sphere
mesh file's node coordinates; a corner at height z is at radius R + z,
X = r cos(lat) cos(lon), Y = r cos(lat) sin(lon), Z = r sin(lat).
"""

from __future__ import annotations

import numpy as np

W0_VERTICES = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)


def gauss_rule(n: int):
    """(points, weights) of the n-point Gauss-Legendre rule on [0, 1]."""
    t, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (t + 1.0), 0.5 * w


def quadrature_xyoz(n_h: int, n_v: int | None = None):
    """points_xy (n_h^2, 2) with y fastest, weights_xy, points_z, weights_z."""
    n_v = n_h if n_v is None else n_v
    px, wx = gauss_rule(n_h)
    pz, wz = gauss_rule(n_v)
    xx, yy = np.meshgrid(px, px, indexing="ij")        # i (x) outer, j (y) inner
    wxy = np.outer(wx, wx).ravel()
    return np.stack([xx.ravel(), yy.ravel()], axis=1), wxy, pz, wz


def w0_basis(points_xy, points_z):
    """basis (8, nqp_h, nqp_v) and reference gradients (3, 8, nqp_h, nqp_v)."""
    x = points_xy[:, 0][:, None]
    y = points_xy[:, 1][:, None]
    z = points_z[None, :]
    one = np.ones_like(x * z)
    f = lambda v, s: s if v else 1.0 - s
    d = lambda v: 1.0 if v else -1.0
    B = np.empty((8,) + one.shape)
    G = np.empty((3, 8) + one.shape)
    for a, (vx, vy, vz) in enumerate(W0_VERTICES):
        fx, fy, fz = f(vx, x) * one, f(vy, y) * one, f(vz, z) * one
        B[a] = fx * fy * fz
        G[0, a] = d(vx) * fy * fz
        G[1, a] = fx * d(vy) * fz
        G[2, a] = fx * fy * d(vz)
    return B, G


def shell_corners(node_lon_deg, node_lat_deg, face_nodes, radius: float, heights):
    """chi (ncell, nlayers, 8, 3): Cartesian corners of every cell-layer, corner
    order of the W0 basis (bottom SW, SE, NE, NW, top SW, SE, NE, NW)."""
    lon = np.deg2rad(np.asarray(node_lon_deg, np.float64))
    lat = np.deg2rad(np.asarray(node_lat_deg, np.float64))
    unit = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=1)
    u = unit[np.asarray(face_nodes) - 1]                 # (ncell, 4, 3)
    h = np.asarray(heights, np.float64)
    nl = h.size - 1
    r_bot = (radius + h[:-1])[None, :, None, None]
    r_top = (radius + h[1:])[None, :, None, None]
    bot = r_bot * u[:, None, :, :]
    top = r_top * u[:, None, :, :]
    return np.concatenate([bot, top], axis=2)             # (ncell, nl, 8, 3)


def jacobian(chi, G):
    """jac (..., 3, 3, nqp_h, nqp_v) = sum_df chi_i(df) dphi_df/dxi_dir; det (..., nqp_h, nqp_v)."""
    jac = np.einsum("...di,edhv->...iehv", chi, G)       # [i (coord), dir]
    j = lambda a, b: jac[..., a, b, :, :]
    det = (j(0, 0) * (j(1, 1) * j(2, 2) - j(1, 2) * j(2, 1))
           - j(0, 1) * (j(1, 0) * j(2, 2) - j(1, 2) * j(2, 0))
           + j(0, 2) * (j(1, 0) * j(2, 1) - j(1, 1) * j(2, 0)))
    return jac, det


def w0_mass_stiffness(chi, n_q: int = 3):
    """Local W0 mass and stiffness matrices, each (..., 8, 8), for corners chi (..., 8, 3)."""
    pxy, wxy, pz, wz = quadrature_xyoz(n_q)
    B, G = w0_basis(pxy, pz)
    jac, det = jacobian(chi, G)
    w = wxy[:, None] * wz[None, :]                       # (nqp_h, nqp_v)
    wd = w * det                                         # w_h w_v detJ
    mass = np.einsum("ahv,bhv,...hv->...ab", B, B, wd)
    # physical gradients: J^T g = ghat  ->  g = J^-T ghat
    Jt = np.moveaxis(jac, (-4, -3), (-2, -1))            # (..., h, v, 3(i), 3(dir))
    JtT = np.swapaxes(Jt, -1, -2)                        # [dir, i]
    ghat = np.moveaxis(G, (0, 1), (-2, -1))              # (h, v, 3(dir), 8)
    g = np.linalg.solve(JtT, np.broadcast_to(ghat, JtT.shape[:-2] + (3, 8)))   # (..., h, v, 3, 8)
    stiff = np.einsum("...hvia,...hvib,...hv->...ab", g, g, np.abs(wd))
    return mass, stiff


def to_local_stencil(A):
    """(ncell, nlayers, ndf_to, ndf_from) -> local_stencil layout (ndf_from, ndf_to, ncell_3d)."""
    nc, nl, nt, nf = A.shape
    return np.ascontiguousarray(A.reshape(nc * nl, nt, nf).transpose(2, 1, 0))


def w0_chi_fields(fs, chi):
    """The coordinate field as LFRic holds it: chi1, chi2, chi3 on the
    lowest-order W0 space ``fs`` ((3, undf) FP64), from the cell-layer
    corners chi (ncell, nlayers, 8, 3) in the W0 local-dof order (corner j of
    cell c at layer k is dof map(j, c) + k; shared vertices receive the same
    values from every cell)."""
    nc, nl = chi.shape[0], chi.shape[1]
    out = np.zeros((3, fs.undf))
    dm = fs.dofmap.astype(np.int64) - 1                   # (ncell, 8)
    d = dm[:, None, :] + np.arange(nl)[None, :, None]     # (ncell, nl, 8)
    for i in range(3):
        out[i, d.ravel()] = chi[..., i].ravel()
    return out
