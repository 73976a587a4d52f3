"""Cubed-sphere topology against the reference's own UGRID mesh files.

The fixtures are the face-face and face-node tables of the reference's
mesh_C4/C8/C12/C16/C24/C32 files (tests/golden/make_fixtures.py).  Face links
must match exactly; vertex tables must match after the local mesh's
first-encounter renumbering (infrastructure/source/mesh/local_mesh_mod.f90:348-372).
"""
import numpy as np
import pytest

from lfric_core_amd.mesh import (E, S, canonical_vertex_ids, cubed_sphere, extrusion_eta,
                                 unit_test_mesh)

from conftest import GOLDEN

MESHES = np.load(GOLDEN / "meshes.npz")


@pytest.mark.parametrize("n", [4, 8, 12, 16, 24, 32])
def test_cubed_sphere_matches_reference_mesh_files(n):
    m = cubed_sphere(n)
    links = MESHES[f"C{n}_face_links"]
    nodes = MESHES[f"C{n}_face_nodes"]
    np.testing.assert_array_equal(m.cell_next, links)
    np.testing.assert_array_equal(m.vert_on_cell, canonical_vertex_ids(nodes.astype(np.int64)))
    assert m.nverts == 6 * n * n + 2
    assert m.nedges == 12 * n * n


@pytest.mark.parametrize("n", [2, 3, 5, 48])
def test_cubed_sphere_invariants(n):
    m = cubed_sphere(n)
    cn = m.cell_next
    ids = np.arange(1, m.ncells + 1)
    # every neighbour relation is symmetric
    for s in range(4):
        back = (cn[cn[:, s] - 1] == ids[:, None]).sum(axis=1)
        assert np.all(back == 1)
    # Euler characteristic of the sphere: V - E + F = 2
    assert m.nverts - m.nedges + m.ncells == 2
    # every vertex is shared by 3 (cube corners, 8 of them) or 4 cells
    counts = np.bincount(m.vert_on_cell.ravel())[1:]
    assert set(np.unique(counts)) == {3, 4}
    assert (counts == 3).sum() == 8


def test_unit_test_mesh_is_biperiodic_3x3():
    m = unit_test_mesh()
    assert m.ncells == 9 and m.nverts == 9
    ids = np.arange(1, 10)
    for s in range(4):
        assert np.all((m.cell_next[m.cell_next[:, s] - 1] == ids[:, None]).sum(axis=1) == 1)


def test_extrusion():
    np.testing.assert_allclose(extrusion_eta(4), [0, .25, .5, .75, 1])
    np.testing.assert_allclose(extrusion_eta(4, "quadratic"), [0, .0625, .25, .5625, 1])


@pytest.mark.parametrize("coarse,fine", [(16, 32), (8, 16), (4, 8), (4, 12)])
def test_intergrid_map_matches_reference_multigrid_file(coarse, fine):
    from lfric_core_amd.mesh import coarse_of_fine, intergrid_map
    g = np.load(GOLDEN / "meshes.npz")
    np.testing.assert_array_equal(intergrid_map(coarse, fine), g[f"C{coarse}_C{fine}_map"])
    np.testing.assert_array_equal(coarse_of_fine(coarse, fine), g[f"C{fine}_C{coarse}_map"].ravel())


def test_intergrid_map_children_are_nested():
    """Children of a coarse cell form a connected 2x2 block: x-neighbours are east of each other."""
    from lfric_core_amd.mesh import intergrid_map
    fine = cubed_sphere(16)
    m = intergrid_map(8, 16)
    for c in range(m.shape[0]):
        for y in range(2):
            assert fine.cell_next[m[c, y, 0] - 1, E] == m[c, y, 1]
        for x in range(2):
            assert fine.cell_next[m[c, 0, x] - 1, S] == m[c, 1, x]
