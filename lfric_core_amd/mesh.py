"""Horizontal meshes and vertical extrusion for the operator path.

Only what the LMA/CMA kernels consume is built here: cell order, the
face-neighbour table ``cell_next(4, ncell)`` in W,S,E,N order and the corner
table ``vert_on_cell(4, ncell)`` in SW,SE,NE,NW order, both 1-based as in the
reference.  Vertex ids carry no meaning beyond equality (the reference's local
mesh renumbers them by first encounter,
infrastructure/source/mesh/local_mesh_mod.f90:262-545), so ``vert_on_cell``
is returned in that canonical first-encounter numbering.

Cubed sphere
------------
Cell numbering and panel layout follow the reference generator
(mesh_tools/source/support/gencube_ps_mod.F90):

* panels 1-4 run eastward round the equator, 5 is the north cap, 6 the south
  cap (:385-398 diagram and ``panel_next`` table);
* on each panel cells are numbered row-major from the north-west corner
  (:402-409, ``get_panel_edge_cell_ids`` :1943-1973);
* corners of a cell are stored SW, SE, NE, NW (reference_element_mod.F90:346-349
  aliases used at gencube_ps_mod.F90:60-64);
* panels 3, 4 and 5 have their per-cell tables cyclically shifted by
  ``PANEL_ROTATIONS = (0, 0, 1, 1, -1, 0)`` (:69, ``orient_lfric`` :1561-1602).

The adjacency itself is derived here from geometry rather than from the
reference's per-panel edge tables: every panel is laid on the surface of the
integer cube [0, N]^3 with its east and south axes, so two cells are
neighbours exactly when they share two lattice corners.  The embedding below
was chosen so that it reproduces the reference's panel_next table and its
edge reversals; tests compare the result with the reference's own mesh files.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

W, S, E, N = 0, 1, 2, 3          # 0-based slots of cell_next (reference W=1..N=4)
SW, SE, NE, NW = 0, 1, 2, 3      # 0-based slots of vert_on_cell

PANEL_ROTATIONS = (0, 0, 1, 1, -1, 0)


@dataclass
class Mesh2D:
    """A 2D local mesh as the reference's local_mesh_type presents it (serial)."""

    name: str
    ncells: int
    cell_next: np.ndarray          # (ncells, 4) int32, 1-based, 0 = no neighbour
    vert_on_cell: np.ndarray       # (ncells, 4) int32, 1-based first-encounter ids
    npanels: int = 1
    edge_cells: int = 0            # cubed sphere: cells along a panel edge
    global_cell_id: np.ndarray = field(default=None)  # (ncells,) 1-based

    def __post_init__(self):
        if self.global_cell_id is None:
            self.global_cell_id = np.arange(1, self.ncells + 1, dtype=np.int32)

    @property
    def nverts(self) -> int:
        return int(self.vert_on_cell.max())

    @property
    def nedges(self) -> int:
        # Closed quadrilateral surfaces: every edge is shared by two cells.
        nbr = (self.cell_next > 0).sum()
        return int(nbr // 2 + (self.cell_next == 0).sum())


def canonical_vertex_ids(vert_keys: np.ndarray) -> np.ndarray:
    """Renumber vertex keys 1.. by first encounter in (cell, slot) order.

    Same rule as local_mesh_mod.f90:348-372 (a vertex gets the next local id
    the first time a cell in local order lists it).
    """
    flat = vert_keys.reshape(-1)
    _, first, inv = np.unique(flat, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return (rank[inv] + 1).reshape(vert_keys.shape).astype(np.int32)


def _panel_point(p: int, a: np.ndarray, b: np.ndarray, n: int):
    """Lattice point of panel ``p`` (1..6) at east offset ``a``, south offset ``b``.

    Panel normals +x, +y, -x, -y, +z, -z; every panel frame is right handed
    (east x north = outward normal).
    """
    z0 = np.zeros_like(a)
    if p == 1:
        return (z0 + n, a, n - b)
    if p == 2:
        return (n - a, z0 + n, n - b)
    if p == 3:
        return (z0, n - a, n - b)
    if p == 4:
        return (a, z0, n - b)
    if p == 5:
        return (n - a, b, z0 + n)
    if p == 6:
        return (n - a, n - b, z0)
    raise ValueError(p)


def cubed_sphere(edge_cells: int) -> Mesh2D:
    """Topology of the C<edge_cells> cubed sphere in the reference's numbering."""
    n = int(edge_cells)
    if n < 2:
        raise ValueError("cubed sphere needs at least 2 cells per panel edge")
    cpp = n * n
    ncells = 6 * cpp
    r, c = np.divmod(np.arange(cpp, dtype=np.int64), n)   # row from north, col from west
    keys = np.empty((ncells, 4), dtype=np.int64)
    m = n + 1

    def key(pt):
        x, y, z = pt
        return (x * m + y) * m + z

    for p in range(1, 7):
        sl = slice((p - 1) * cpp, p * cpp)
        keys[sl, SW] = key(_panel_point(p, c, r + 1, n))
        keys[sl, SE] = key(_panel_point(p, c + 1, r + 1, n))
        keys[sl, NE] = key(_panel_point(p, c + 1, r, n))
        keys[sl, NW] = key(_panel_point(p, c, r, n))

    # Side k of a cell joins corners (side_corners[k][0], side_corners[k][1]).
    side_corners = {W: (NW, SW), S: (SW, SE), E: (SE, NE), N: (NE, NW)}
    ek = np.empty((ncells, 4), dtype=np.int64)
    big = np.int64(m) ** 3
    for s_, (c0, c1) in side_corners.items():
        lo = np.minimum(keys[:, c0], keys[:, c1])
        hi = np.maximum(keys[:, c0], keys[:, c1])
        ek[:, s_] = lo * big + hi
    flat = ek.reshape(-1)
    order = np.argsort(flat, kind="stable")
    sorted_keys = flat[order]
    if not np.all(sorted_keys[0::2] == sorted_keys[1::2]):
        raise RuntimeError("cubed-sphere edge not shared by exactly two cells")
    other = np.empty_like(order)
    other[order[0::2]] = order[1::2]
    other[order[1::2]] = order[0::2]
    cell_next = (other // 4 + 1).reshape(ncells, 4).astype(np.int32)

    # Per-panel cyclic shift of the slot order (Fortran cshift along dim 1:
    # result(i) = array(i + shift)).
    for p, rot in enumerate(PANEL_ROTATIONS, start=1):
        if rot == 0:
            continue
        sl = slice((p - 1) * cpp, p * cpp)
        cell_next[sl] = np.roll(cell_next[sl], -rot, axis=1)
        keys[sl] = np.roll(keys[sl], -rot, axis=1)

    return Mesh2D(name=f"C{n}", ncells=ncells, cell_next=cell_next,
                  vert_on_cell=canonical_vertex_ids(keys), npanels=6,
                  edge_cells=n)


def unit_test_mesh() -> Mesh2D:
    """The 3x3 biperiodic mesh of local_mesh_mod.f90:1269-1410 (initialise_unit_test)."""
    cell_next = np.array([
        [3, 7, 2, 4], [1, 8, 3, 5], [2, 9, 1, 6],
        [6, 1, 5, 7], [4, 2, 6, 8], [5, 3, 4, 9],
        [9, 4, 8, 1], [7, 5, 9, 2], [8, 6, 7, 3]], dtype=np.int32)
    vert_on_cell = np.array([
        [1, 2, 3, 4], [2, 5, 6, 3], [5, 1, 4, 6],
        [4, 3, 7, 8], [3, 6, 9, 7], [6, 4, 8, 9],
        [8, 7, 2, 1], [7, 9, 5, 2], [9, 8, 1, 5]], dtype=np.int32)
    return Mesh2D(name="unit_test", ncells=9, cell_next=cell_next,
                  vert_on_cell=vert_on_cell, npanels=1)


def from_ugrid_arrays(face_links: np.ndarray, face_nodes: np.ndarray,
                      npanels: int, name: str = "ugrid") -> Mesh2D:
    """Serial local mesh from UGRID face tables (1-based, fill value < 0 -> 0)."""
    cn = np.where(face_links > 0, face_links, 0).astype(np.int32)
    ncells = cn.shape[0]
    edge_cells = int(round(np.sqrt(ncells / 6))) if npanels == 6 else 0
    return Mesh2D(name=name, ncells=ncells, cell_next=cn,
                  vert_on_cell=canonical_vertex_ids(face_nodes.astype(np.int64)),
                  npanels=npanels, edge_cells=edge_cells)


def extrusion_eta(nlayers: int, kind: str = "uniform") -> np.ndarray:
    """Nondimensional level heights eta(0..nlayers).

    uniform: k/nl (infrastructure/source/mesh/extrusion_mod.f90:292-302);
    quadratic: (k/nl)**2 (:341-351).
    """
    k = np.arange(nlayers + 1, dtype=np.float64) / float(nlayers)
    if kind == "uniform":
        return k
    if kind == "quadratic":
        return k ** 2
    raise ValueError(kind)


def intergrid_map(coarse_n: int, fine_n: int) -> np.ndarray:
    """Fine cells under each coarse cell of two nested cubed spheres.

    Returns map[coarse_cell-1, y, x] = fine cell id (1-based), the layout of
    the reference's mesh-mesh connectivity (cell_map(x, y, cell) in Fortran
    order), following calc_global_cell_map
    (mesh_tools/source/support/calc_global_cell_map_mod.F90:134-258): on each
    panel the fine cells of a coarse cell are the ef x ef block below it in
    panel row-major numbering, turned for rotated panels (PANEL_ROTATIONS) so
    that (x, y) follow the coarse cell's own W->E, N->S directions.
    """
    if fine_n % coarse_n:
        raise ValueError("fine mesh must refine the coarse mesh by an integer factor")
    ef = fine_n // coarse_n
    ccpp, fcpp = coarse_n * coarse_n, fine_n * fine_n
    out = np.empty((6 * ccpp, ef, ef), dtype=np.int32)
    ii, jj = np.meshgrid(np.arange(ef), np.arange(ef), indexing="xy")   # [y, x] grids
    for p in range(6):
        rot = PANEL_ROTATIONS[p]
        for j in range(coarse_n):          # coarse row (y)
            for i in range(coarse_n):      # coarse column (x)
                # tmp(a, b) = fine cell at column i*ef + a, row j*ef + b (0-based)
                def tmp(a, b):
                    return p * fcpp + (j * ef + b) * fine_n + (i * ef + a) + 1
                c = p * ccpp + j * coarse_n + i
                if rot == 0:
                    out[c] = tmp(ii, jj)
                elif rot == 1:
                    # map(ifine, jfine) = tmp(jfine, ef+1-ifine)
                    out[c] = tmp(jj, ef - 1 - ii)
                else:
                    # map(ifine, jfine) = tmp(ef+1-jfine, ifine)
                    out[c] = tmp(ef - 1 - jj, ii)
    return out


def coarse_of_fine(coarse_n: int, fine_n: int) -> np.ndarray:
    """parent[fine_cell-1] = coarse cell id (1-based)."""
    m = intergrid_map(coarse_n, fine_n)
    parent = np.empty(6 * fine_n * fine_n, dtype=np.int32)
    parent[m.reshape(m.shape[0], -1) - 1] = np.arange(1, m.shape[0] + 1)[:, None]
    return parent
