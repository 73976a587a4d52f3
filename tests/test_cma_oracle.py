"""CMA operator geometry (lfric_core_amd/cma.py) and the oracle's
columnwise_op_asm / columnwise_op_app, on the CPU.

Pinned by: the reference's own test assertions for a Wtheta operator
(alpha = 1, gamma_m = 1, bandwidth 3:
components/science/unit-test/kernel/columnwise_operators/
sci_columnwise_op_appinv_kernel_mod_test.pf:107-110); the flang-built reference
kernels (bit-for-bit); and the meaning of the operator: assembling an LMA
operator into banded form and applying it must equal the LMA apply itself."""
import numpy as np
import pytest

import oracle
from lfric_core_amd import function_space as fsm
from lfric_core_amd.cma import cma_geometry
from lfric_core_amd.configs import normal_lma
from lfric_core_amd.mesh import cubed_sphere, unit_test_mesh
from lfric_core_amd.rng import Stream

PAIRS = [("W3", "W3"), ("WTHETA", "WTHETA"), ("WTHETA", "W3"), ("W3", "WTHETA"), ("W2V", "W2V"),
         ("W3", "W2V")]


def test_geometry_reference_assertions():
    m = unit_test_mesh()
    wt = fsm.build_dofmap(m, "WTHETA", 3)
    g = cma_geometry(wt, wt)
    assert (g.alpha, g.gamma_m, g.bandwidth) == (1, 1, 3)
    assert (g.nrow, g.ncol) == (4, 4)
    w3 = fsm.build_dofmap(m, "W3", 3)
    g = cma_geometry(wt, w3)            # to Wtheta, from W3
    assert (g.alpha, g.beta, g.gamma_m, g.gamma_p, g.bandwidth, g.nrow, g.ncol) == (1, 1, 0, 1, 2, 4, 3)


@pytest.mark.parametrize("to,fr", PAIRS)
def test_indirection_is_a_column_permutation(to, fr):
    m = cubed_sphere(3)
    g = cma_geometry(fsm.build_dofmap(m, to, 5), fsm.build_dofmap(m, fr, 5))
    assert sorted(g.indirection_to) == list(range(1, g.nrow + 1))
    assert sorted(g.indirection_from) == list(range(1, g.ncol + 1))


def _problem(to, fr, nl=5, edge=4, seed=9, dtype=np.float32):
    m = cubed_sphere(edge)
    ft = fsm.build_dofmap(m, to, nl)
    ff = ft if to == fr else fsm.build_dofmap(m, fr, nl)
    st = Stream(seed)
    M = normal_lma(st, m.ncells * nl, ft.ndf, ff.ndf).astype(dtype)
    x = st.uniform(ff.undf).astype(dtype)
    return m, ft, ff, M, x


@pytest.mark.parametrize("to,fr", PAIRS)
def test_asm_and_app_match_reference_bitwise(to, fr):
    if oracle.ref() is None:
        pytest.skip("reference build (oracle/_ref) not available")
    m, ft, ff, M, x = _problem(to, fr)
    g = cma_geometry(ft, ff)
    cm = oracle.cma_asm(M, g, m.ncells, ft.nlayers)
    cm_ref = oracle.cma_asm(M, g, m.ncells, ft.nlayers, use_ref=True)
    np.testing.assert_array_equal(cm, cm_ref)
    y = oracle.cma_app(np.zeros(ft.undf, np.float32), x, cm, g, ft.dofmap, ff.dofmap)
    y_ref = oracle.cma_app(np.zeros(ft.undf, np.float32), x, cm, g, ft.dofmap, ff.dofmap, use_ref=True)
    np.testing.assert_array_equal(y, y_ref)


@pytest.mark.parametrize("to,fr", PAIRS)
def test_banded_apply_equals_lma_apply(to, fr):
    m, ft, ff, M, x = _problem(to, fr, dtype=np.float64)
    g = cma_geometry(ft, ff)
    cm = oracle.cma_asm(M, g, m.ncells, ft.nlayers)
    y = oracle.cma_app(np.zeros(ft.undf), x, cm, g, ft.dofmap, ff.dofmap)
    col = fsm.colour_cells(m)
    y_lma = oracle.invoke_coloured(0, col, ft.nlayers, np.zeros(ft.undf), x, M, ft.dofmap, ff.dofmap)
    np.testing.assert_allclose(y, y_lma, rtol=1e-12, atol=1e-12 * np.abs(y_lma).max())
