"""Local mass matrices by quadrature, restated in numpy (TEST INFRASTRUCTURE
ONLY: the checker of lfr_compute_mass_matrix and
lfr_compute_derham_matrices).

* scalar spaces: compute_mass_matrix_w_scalar_code
  (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w_scalar_mod.F90:150-174)
  mm(ik, df, df2) = sum_{qp_v} sum_{qp_h} w_h w_v phi_df phi_df2 detJ
* W2 with the Piola map: compute_mass_matrix_w2_code
  (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w2_mod.F90:131-162)
  mm(ik, df, df2) = sum w_h w_v (J v_df) . (J v_df2) / detJ
* the lowest-order de Rham complex: compute_derham_matrices_code
  (components/science/source/kernel/fem/sci_compute_derham_matrices_kernel_mod.F90:241-352):
  the six mass matrices, W1 with the covariant map Jt = transpose(J^-1)
  (J^-1 as pointwise_coordinate_jacobian_inverse,
  components/science/source/kernel/geometry/sci_coordinate_jacobian_mod.F90:1140-1161),
  and grad = sum wt (Jt w1_df).(Jt grad phi0_df2) detJ,
  curl = sum wt (J w2_df).(J curl w1_df2) / detJ,
  div = sum wt w3_df div w2_df2 (no metric term)

Mass matrices for df >= df2, then mirrored; the Jacobian of the chi map is
jac(i, dir) = sum_df chi_i(df) dphi_df/dxi_dir with its cofactor determinant
(components/science/source/kernel/geometry/sci_coordinate_jacobian_mod.F90:330-345,430-440).
The sums run in the reference's order (qp_v outer, qp_h inner), one
cell-layer per row of the vectorised arrays.  The basis tables are inputs, as
the PSy layer passes them to the kernels.

Pinned (tests/test_mass_oracle.py) by: the W0 mass of the independent host
assembly (einsum sums) on the C12 shell, the W3 mass = the cell volume, the
Wtheta and W2 masses on axis-aligned boxes in closed form (integrals of the
linear basis products), and the W2 mass on sheared affine cells against
(1/det A) sum_ij (A^T A)_ij I_ij, I from the tables alone; the W1 mass
likewise with det A (A^-1 A^-T); grad = M1 G, curl = M2 C and div = D for
the integer incidence matrices G, C, D of the complex (found from the
tables, D C = 0 and C G = 0) on curved shell cells.  The reference
kernels use generated configuration modules (base_mesh_config_mod,
finite_element_config_mod), so they are not compiled here.
"""

from __future__ import annotations

import numpy as np


def jacobian(chi_e, diff_basis_chi):
    """chi_e (n, 3, ndf_chi) element coordinates; diff_basis_chi
    (3, ndf_chi, nqp_h, nqp_v).  Returns jac (n, 3, 3, nqp_h, nqp_v) and dj
    (n, nqp_h, nqp_v), accumulated over df in order."""
    n, _, ndf = chi_e.shape
    nh, nv = diff_basis_chi.shape[2:]
    jac = np.zeros((n, 3, 3, nh, nv))
    for df in range(ndf):
        for i in range(3):
            for d in range(3):
                jac[:, i, d] = jac[:, i, d] + chi_e[:, i, df, None, None] * diff_basis_chi[d, df][None]
    j = lambda a, b: jac[:, a, b]
    dj = (j(0, 0) * (j(1, 1) * j(2, 2) - j(1, 2) * j(2, 1))
          - j(0, 1) * (j(1, 0) * j(2, 2) - j(1, 2) * j(2, 0))
          + j(0, 2) * (j(1, 0) * j(2, 1) - j(1, 1) * j(2, 0)))
    return jac, dj


def mass_matrix(kind, basis, diff_basis_chi, wqp_h, wqp_v, chi_e):
    """Local mass matrices (n, ndf, ndf), [ik, df, df2] = mm(ik, df, df2).

    kind 0: scalar space, basis (1, ndf, nqp_h, nqp_v); kind 1: W2 (Piola),
    basis (3, ndf, nqp_h, nqp_v).  chi_e (n, 3, ndf_chi)."""
    basis = np.asarray(basis, np.float64)
    ndf, nh, nv = basis.shape[1:]
    jac, dj = jacobian(np.asarray(chi_e, np.float64), np.asarray(diff_basis_chi, np.float64))
    n = dj.shape[0]
    mm = np.zeros((n, ndf, ndf))
    if kind == 1:
        # J v per dof: jv[:, df, i] = sum_d jac(i, d) v_d(df)
        jv = np.zeros((n, ndf, 3, nh, nv))
        for df in range(ndf):
            for i in range(3):
                s = np.zeros((n, nh, nv))
                for d in range(3):
                    s = s + jac[:, i, d] * basis[d, df][None]
                jv[:, df, i] = s
    for df2 in range(ndf):
        for df in range(df2, ndf):
            acc = np.zeros(n)
            for q2 in range(nv):
                for q1 in range(nh):
                    wq = wqp_h[q1] * wqp_v[q2]
                    if kind == 0:
                        v = ((wq * basis[0, df, q1, q2]) * basis[0, df2, q1, q2]) * dj[:, q1, q2]
                    else:
                        dot = np.zeros(n)
                        for i in range(3):
                            dot = dot + jv[:, df, i, q1, q2] * jv[:, df2, i, q1, q2]
                        v = (wq * dot) / dj[:, q1, q2]
                    acc = acc + v
            mm[:, df, df2] = acc
            mm[:, df2, df] = acc
    return mm


def _inverse(jac, dj):
    """jac_inv (n, 3, 3, h, v) as pointwise_coordinate_jacobian_inverse:
    cofactor expressions times 1/dj."""
    j = lambda a, b: jac[:, a, b]
    idj = 1.0 / dj
    inv = np.empty_like(jac)
    inv[:, 0, 0] = (j(1, 1) * j(2, 2) - j(1, 2) * j(2, 1)) * idj
    inv[:, 0, 1] = -(j(0, 1) * j(2, 2) - j(0, 2) * j(2, 1)) * idj
    inv[:, 0, 2] = (j(0, 1) * j(1, 2) - j(0, 2) * j(1, 1)) * idj
    inv[:, 1, 0] = -(j(1, 0) * j(2, 2) - j(1, 2) * j(2, 0)) * idj
    inv[:, 1, 1] = (j(0, 0) * j(2, 2) - j(0, 2) * j(2, 0)) * idj
    inv[:, 1, 2] = -(j(0, 0) * j(1, 2) - j(0, 2) * j(1, 0)) * idj
    inv[:, 2, 0] = (j(1, 0) * j(2, 1) - j(1, 1) * j(2, 0)) * idj
    inv[:, 2, 1] = -(j(0, 0) * j(2, 1) - j(0, 1) * j(2, 0)) * idj
    inv[:, 2, 2] = (j(0, 0) * j(1, 1) - j(0, 1) * j(1, 0)) * idj
    return inv


def _mapped(mat, table):
    """(n, ndf, 3, h, v): matmul(mat, table(:, df)) summed over d in order."""
    n, _, _, nh, nv = mat.shape
    ndf = table.shape[1]
    out = np.zeros((n, ndf, 3, nh, nv))
    for df in range(ndf):
        for i in range(3):
            s = np.zeros((n, nh, nv))
            for d in range(3):
                s = s + mat[:, i, d] * table[d, df][None]
            out[:, df, i] = s
    return out


def form(mapping, L, R, diff_basis_chi, wqp_h, wqp_v, chi_e, symmetric=False):
    """One bilinear form of compute_derham_matrices_code, (n, nrow, ncol):
    mapping "dj" ((wt L) R) detJ, "cov" (wt (Jt L).(Jt R)) detJ, "contra"
    (wt (J L).(J R)) / detJ, "ref" (wt L) R.  Symmetric forms are computed
    for df >= df2 and mirrored; sums qp_v outer, qp_h inner."""
    L = np.asarray(L, np.float64)
    R = np.asarray(R, np.float64)
    jac, dj = jacobian(np.asarray(chi_e, np.float64), np.asarray(diff_basis_chi, np.float64))
    n, nh, nv = dj.shape
    nr, nc = L.shape[1], R.shape[1]
    if mapping in ("cov", "contra"):
        mat = np.swapaxes(_inverse(jac, dj), 1, 2) if mapping == "cov" else jac
        Lm, Rm = _mapped(mat, L), _mapped(mat, R)
    out = np.zeros((n, nr, nc))
    for df2 in range(nc):
        for df in (range(df2, nr) if symmetric else range(nr)):
            acc = np.zeros(n)
            for q2 in range(nv):
                for q1 in range(nh):
                    wt = wqp_h[q1] * wqp_v[q2]
                    if mapping in ("cov", "contra"):
                        dot = np.zeros(n)
                        for i in range(3):
                            dot = dot + Lm[:, df, i, q1, q2] * Rm[:, df2, i, q1, q2]
                        v = (wt * dot) * dj[:, q1, q2] if mapping == "cov" else (wt * dot) / dj[:, q1, q2]
                    elif mapping == "dj":
                        v = ((wt * L[0, df, q1, q2]) * R[0, df2, q1, q2]) * dj[:, q1, q2]
                    else:
                        v = np.full(n, (wt * L[0, df, q1, q2]) * R[0, df2, q1, q2])
                    acc = acc + v
            out[:, df, df2] = acc
            if symmetric:
                out[:, df2, df] = acc
    return out


def derham_matrices(t, chi_e):
    """All ten outputs of compute_derham_matrices_code for the tables t (a
    dict: wqp_h, wqp_v, diff_basis_chi, basis_w0, diff_basis_w0, basis_w1,
    diff_basis_w1, basis_w2, diff_basis_w2, basis_w2b, diff_basis_w2b,
    basis_w3, basis_wt), each (n, ndf_to, ndf_from)."""
    g, wh, wv = t["diff_basis_chi"], t["wqp_h"], t["wqp_v"]
    f = lambda m, a, b, sym=False: form(m, t[a], t[b], g, wh, wv, chi_e, sym)
    return {"mm0": f("dj", "basis_w0", "basis_w0", True),
            "mm1": f("cov", "basis_w1", "basis_w1", True),
            "mm2": f("contra", "basis_w2", "basis_w2", True),
            "mm2b": f("contra", "basis_w2b", "basis_w2b", True),
            "mm3": f("dj", "basis_w3", "basis_w3", True),
            "mmt": f("dj", "basis_wt", "basis_wt", True),
            "grad": f("cov", "basis_w1", "diff_basis_w0"),
            "curl": f("contra", "basis_w2", "diff_basis_w1"),
            "div": f("ref", "basis_w3", "diff_basis_w2"),
            "broken_div": f("ref", "basis_w3", "diff_basis_w2b")}
