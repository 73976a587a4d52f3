"""Vertical tridiagonal solves (tri_solve, operator_tri_solve) in the oracle.

tri_solve_code (components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90)
cannot be compiled here (its module uses fs_continuity_mod, whose log_mod needs
MPI), so it is pinned by what the mathematics fixes: the Thomas recurrence is
the LU solve of the column's tridiagonal system, checked against
scipy.linalg.solve_banded, and it shares its recurrence with
operator_tri_solve_code, which is checked bit for bit against the flang-built
reference kernel and against a dense solve of the assembled column operator.
"""

import numpy as np
import pytest
from scipy.linalg import solve_banded

import oracle
from lfric_core_amd.function_space import build_dofmap
from lfric_core_amd.mesh import cubed_sphere
from lfric_core_amd.rng import Stream


def _w3(nc=3, nl=10):
    return build_dofmap(cubed_sphere(nc), "W3", nl)


def _bands(st, n, dtype):
    t0 = (4.0 + st.uniform(n)).astype(dtype)          # diagonally dominant
    tp = (0.5 * st.uniform(n)).astype(dtype)
    tm = (0.5 * st.uniform(n)).astype(dtype)
    return t0, tp, tm


def test_tri_solve_is_the_banded_lu_solve():
    fs = _w3()
    nl = fs.nlayers
    st = Stream(41)
    t0, tp, tm = _bands(st, fs.undf, np.float64)
    x = st.uniform(fs.undf)
    y = np.full(fs.undf, np.nan)
    oracle.tri_solve(y, x, t0, tp, tm, fs.dofmap, nl)
    for c in range(fs.mesh.ncells):
        s = fs.dofmap[c, 0] - 1
        sl = slice(s, s + nl)
        ab = np.zeros((3, nl))
        ab[0, 1:] = tp[sl][:-1]        # super-diagonal: row k couples k+1 through tri_plus(k)
        ab[1] = t0[sl]
        ab[2, :-1] = tm[sl][1:]        # sub-diagonal: row k couples k-1 through tri_minus(k)
        ref = solve_banded((1, 1), ab, x[sl])
        np.testing.assert_allclose(y[sl], ref, rtol=1e-13, atol=1e-14)
    # residual of the assembled system, all columns at once
    r = t0 * y - x
    for c in range(fs.mesh.ncells):
        s = fs.dofmap[c, 0] - 1
        r[s:s + nl - 1] += tp[s:s + nl - 1] * y[s + 1:s + nl]
        r[s + 1:s + nl] += tm[s + 1:s + nl] * y[s:s + nl - 1]
    assert np.abs(r).max() < 1e-13


def test_tri_solve_fp32_tracks_fp64():
    fs = _w3(2, 7)
    st = Stream(42)
    t0, tp, tm = _bands(st, fs.undf, np.float64)
    x = st.uniform(fs.undf)
    y64 = np.zeros(fs.undf)
    y32 = np.zeros(fs.undf, np.float32)
    oracle.tri_solve(y64, x, t0, tp, tm, fs.dofmap, fs.nlayers)
    oracle.tri_solve(y32, x.astype(np.float32), t0.astype(np.float32), tp.astype(np.float32),
                     tm.astype(np.float32), fs.dofmap, fs.nlayers)
    np.testing.assert_allclose(y32, y64, rtol=2e-5, atol=2e-6)


def _wtheta_problem(nc=3, nl=8, seed=43):
    fs = build_dofmap(cubed_sphere(nc), "WTHETA", nl)
    assert fs.ndf == 2
    st = Stream(seed)
    n3 = fs.mesh.ncells * nl
    # local 2x2 mass-like matrices (SPD, diagonally dominant) stored (ndf, ndf, n3)
    a = 2.0 + st.uniform(n3)
    b = 2.0 + st.uniform(n3)
    o1 = 0.5 * st.uniform(n3)
    o2 = 0.5 * st.uniform(n3)
    m = np.empty((2, 2, n3))
    m[0, 0] = a        # (ik, 1, 1)
    m[0, 1] = o1       # (ik, 2, 1): row 2, column 1
    m[1, 0] = o2       # (ik, 1, 2): row 1, column 2
    m[1, 1] = b        # (ik, 2, 2)
    x = st.uniform(fs.undf)
    return fs, m, x


def test_operator_tri_solve_matches_dense_column_solve():
    fs, m, x = _wtheta_problem()
    nl = fs.nlayers
    y = np.full(fs.undf, np.nan)
    oracle.operator_tri_solve(y, x, m, fs.dofmap, nl)
    n3 = fs.mesh.ncells * nl
    flat = m.reshape(4, n3)    # C order (j, i, ik) = Fortran (ik, i, j): index i + 2 j
    for c in range(fs.mesh.ncells):
        s = fs.dofmap[c, 0] - 1
        A = np.zeros((nl + 1, nl + 1))
        for k in range(nl):
            ik = c * nl + k
            # assemble the local 2x2 (rows: bottom, top dof of layer k)
            A[k, k] += flat[0, ik]
            A[k, k + 1] += flat[2, ik]
            A[k + 1, k] += flat[1, ik]
            A[k + 1, k + 1] += flat[3, ik]
        ref = np.linalg.solve(A, x[s:s + nl + 1])
        np.testing.assert_allclose(y[s:s + nl + 1], ref, rtol=1e-12, atol=1e-13)


@pytest.mark.skipif(oracle.ref() is None, reason="flang-built reference kernels not present")
def test_operator_tri_solve_bitexact_vs_reference_kernel():
    fs, m, x = _wtheta_problem(4, 11, seed=44)
    y_o = np.zeros(fs.undf)
    y_r = np.zeros(fs.undf)
    oracle.operator_tri_solve(y_o, x, m, fs.dofmap, fs.nlayers)
    oracle.operator_tri_solve(y_r, x, m, fs.dofmap, fs.nlayers, use_ref=True)
    np.testing.assert_array_equal(y_o, y_r)
