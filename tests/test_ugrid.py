"""UGRID mesh files (lfric_core_amd.ugrid) against the reference's own mesh
files (tests/golden/ugrid/: copies of infrastructure/unit-test/data/*.nc) and
the fixtures scipy's NetCDF reader wrote from them (tests/golden/meshes.npz):
the NetCDF-3 parser, the mesh attributes and connectivities read_mesh
returns (ncdf_quad_mod.F90:2023-2539), the intergrid maps of read_map
(:2557-2683), and the dofmaps built from a file mesh against those of the
built-in generator."""
import struct

import numpy as np
import pytest

from conftest import ROOT
from lfric_core_amd import ugrid
from lfric_core_amd.function_space import build_dofmap
from lfric_core_amd.mesh import cubed_sphere, intergrid_map

DATA = ROOT / "tests" / "golden" / "ugrid"
GOLD = np.load(ROOT / "tests" / "golden" / "meshes.npz")


def test_c4_connectivity_matches_the_fixture():
    m = ugrid.read_mesh(DATA / "mesh_C4.nc")
    assert m.name == "unit_test" and m.nfaces == 96 and m.npanels == 6
    np.testing.assert_array_equal(m.face_face_connectivity, GOLD["C4_face_links"])
    np.testing.assert_array_equal(m.face_node_connectivity, GOLD["C4_face_nodes"])
    assert (m.geometry, m.topology, m.coord_sys) == ("spherical", "periodic", "ll")
    assert m.coord_units == ("degrees_east", "degrees_north")
    assert m.node_coordinates.shape == (2, 98) and m.face_coordinates.shape == (2, 96)
    # every edge joins two nodes of the faces it bounds
    fe, en, fn = m.face_edge_connectivity, m.edge_node_connectivity, m.face_node_connectivity
    for f in range(m.nfaces):
        for e in fe[f]:
            assert set(en[e - 1]) <= set(fn[f])


def test_file_mesh_is_the_generator_mesh():
    """read_mesh -> Mesh2D gives the generator's cell order, W,S,E,N
    neighbours and SW,SE,NE,NW corners, so dofmaps agree entry for entry."""
    m = ugrid.load_mesh(DATA / "mesh_C4.nc")
    g = cubed_sphere(4)
    np.testing.assert_array_equal(m.cell_next, g.cell_next)
    np.testing.assert_array_equal(m.vert_on_cell, g.vert_on_cell)
    for space in ("W0", "W1", "W2", "W3", "WTHETA"):
        np.testing.assert_array_equal(build_dofmap(m, space, 3).dofmap, build_dofmap(g, space, 3).dofmap,
                                      err_msg=space)


@pytest.mark.parametrize("fname,periodic,topology", [("mesh_planar.nc", (False, False), "non_periodic"),
                                                     ("mesh_trench_x.nc", (True, False), "channel")])
def test_planar_meshes(fname, periodic, topology):
    m = ugrid.read_mesh(DATA / fname)
    assert m.geometry == "planar" and m.topology == topology and m.periodic_xy == periodic
    assert m.coord_units == ("m", "m") and m.npanels == 1 and m.void_cell == -9999
    mesh = m.to_mesh2d()
    links = m.face_face_connectivity
    # void neighbours become 0 in cell_next; real ones are kept
    np.testing.assert_array_equal(mesh.cell_next == 0, links == m.void_cell)
    np.testing.assert_array_equal(mesh.cell_next[links > 0], links[links > 0])
    # a non-periodic 8x8 plane has 4*8 boundary sides; the x-channel has 2*8
    assert int((links == m.void_cell).sum()) == (32 if topology == "non_periodic" else 16)
    # neighbour relation is symmetric
    for c in range(m.nfaces):
        for nb in links[c]:
            if nb > 0:
                assert c + 1 in links[nb - 1]


def test_multi_mesh_file_and_maps():
    f = DATA / "dev_LAM_locals_20-24.nc"
    nc = ugrid.read_netcdf(f)
    assert ugrid.mesh_names(nc) == ["SQ20", "SQ20-lbc", "SQ10"]
    sq20 = ugrid.read_mesh(f, "SQ20")
    assert sq20.maps_to == ["SQ10"] and sq20.periodic_xy == (False, False)
    assert sq20.attrs["max_stencil_depth"] == 1
    np.testing.assert_array_equal(sq20.attrs["north_pole"], [0.0, 90.0])
    fine2coarse = ugrid.read_map(f, "SQ20", "SQ10")
    coarse2fine = ugrid.read_map(f, "SQ10", "SQ20")
    assert fine2coarse.shape[1:] == (1, 1) and coarse2fine.shape[1:] == (2, 2)
    # every fine cell the coarse map names maps back to that coarse cell
    for c in range(coarse2fine.shape[0]):
        for fc in coarse2fine[c].ravel():
            if 0 < fc <= fine2coarse.shape[0]:
                assert fine2coarse[fc - 1, 0, 0] == c + 1
    with pytest.raises(KeyError):
        ugrid.read_mesh(f, "SQ40")
    with pytest.raises(KeyError):
        ugrid.read_map(f, "SQ10", "SQ40")


def test_intergrid_maps_against_the_fixture_and_generator():
    # the multigrid file's maps were written to the fixture by scipy; the
    # generator's intergrid_map reproduces them
    np.testing.assert_array_equal(intergrid_map(4, 8), GOLD["C4_C8_map"])
    np.testing.assert_array_equal(intergrid_map(4, 12), GOLD["C4_C12_map"])


def _write_classic(path, version=1):
    """A small NetCDF-3 file written by hand from the published format: one
    fixed and two record variables, text and numeric attributes."""
    def name(s):
        b = s.encode()
        return struct.pack(">I", len(b)) + b + b"\0" * ((4 - len(b) % 4) % 4)
    off = ">I" if version == 1 else ">Q"
    dims = [("t", 0), ("n", 3)]
    hdr = b"CDF" + bytes([version]) + struct.pack(">I", 2)          # numrecs = 2
    hdr += struct.pack(">II", 0x0A, len(dims)) + b"".join(name(n) + struct.pack(">I", l) for n, l in dims)
    hdr += struct.pack(">II", 0x0C, 1) + name("title") + struct.pack(">II", 2, 5) + b"hello\0\0\0"
    # variables: fixed a(n) int32; records b(t, n) float64 and c(t) int16
    specs = [("a", [1], 4, 12), ("b", [0, 1], 6, 24), ("c", [0], 3, 4)]
    body_len = 0
    var_blocks = []
    for nm, ids, typ, vsize in specs:
        vb = name(nm) + struct.pack(">I", len(ids)) + b"".join(struct.pack(">I", i) for i in ids)
        vb += struct.pack(">II", 0x0C, 1) + name("units") + struct.pack(">II", 2, 1) + b"m\0\0\0"
        vb += struct.pack(">II", typ, vsize)
        var_blocks.append(vb)
    var_hdr = struct.pack(">II", 0x0B, len(specs))
    hdr_len = len(hdr) + len(var_hdr) + sum(len(v) + struct.calcsize(off) for v in var_blocks)
    a_begin = hdr_len
    rec_begin = a_begin + 12
    begins = [a_begin, rec_begin, rec_begin + 24]
    hdr += var_hdr + b"".join(v + struct.pack(off, bgn) for v, bgn in zip(var_blocks, begins))
    assert len(hdr) == hdr_len
    a = np.array([7, -8, 9], ">i4").tobytes()
    recs = b""
    for r in range(2):
        recs += np.array([r + 0.5, r + 1.5, r + 2.5], ">f8").tobytes()
        recs += np.array([10 * r - 3], ">i2").tobytes() + b"\0\0"      # padded to 4
    path.write_bytes(hdr + a + recs)


@pytest.mark.parametrize("version", [1, 2])
def test_netcdf_parser_on_a_hand_written_file(tmp_path, version):
    p = tmp_path / "t.nc"
    _write_classic(p, version)
    nc = ugrid.read_netcdf(p)
    assert nc.dims == {"t": 0, "n": 3} and nc.attrs["title"] == "hello"
    np.testing.assert_array_equal(nc.variables["a"].data, [7, -8, 9])
    np.testing.assert_array_equal(nc.variables["b"].data, [[0.5, 1.5, 2.5], [1.5, 2.5, 3.5]])
    np.testing.assert_array_equal(nc.variables["c"].data, [-3, 7])
    assert nc.variables["b"].dims == ("t", "n") and nc.variables["a"].attrs["units"] == "m"


def test_netcdf_parser_rejects_bad_input(tmp_path):
    p = tmp_path / "x.nc"
    p.write_bytes(b"HDF\x89 not netcdf")
    with pytest.raises(ugrid.UgridError):
        ugrid.read_netcdf(p)
    good = (DATA / "mesh_C4.nc").read_bytes()
    p.write_bytes(good[:200])
    with pytest.raises(ugrid.UgridError):
        ugrid.read_netcdf(p)
