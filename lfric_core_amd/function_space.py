"""Function spaces on an extruded mesh: dof counts, dofmaps and colouring.

This is the host-side half of the reference's function_space_type that the
operator path needs: the per-cell dofmap of the bottom cell of each column
(``map(df) + k`` addresses layer k), the number of unique dofs, and the cell
colouring that the PSy layer uses for continuous (GH_INC) outputs.

* ``ndof_counts`` restates ``ndof_setup``
  (infrastructure/source/function_space/function_space_constructor_helper_functions_mod.F90:147-432)
  for the spaces on the path.
* ``build_dofmap`` reproduces ``dofmap_setup`` (same file :1715-2317) for a
  serial mesh in which every cell is owned: dofs are numbered on first
  encounter walking the cells in order and, inside a cell, volume -> side
  faces W,S,E,N -> the bottom/top face pair -> horizontal edges -> vertical
  edges -> vertices.  Bottom and top copies of an entity share one column of
  nl+1 dofs (top = bottom + 1).  The per-cell output order is volume, faces
  1..6, edges 1..12, vertices 1..8.  Shared faces are matched through
  ``cell_next`` and shared vertices/vertical edges through vertex ids, as
  mesh_constructor_helper_functions_mod.F90:266-438 does.  The loop is
  vectorised: an entity's first occurrence in traversal order owns a block
  of ids whose start is an exclusive prefix sum of block sizes.
* ``colour_cells`` restates ``set_colours_cubed_sphere``, the generic greedy
  fallback ``set_colours_generic`` and ``create_colour_map_lookup``
  (infrastructure/source/mesh/mesh_colouring_mod.F90:151-328, :514-650,
  :703-752) and the conflict check ``test_set_colours`` (:765-837).  The
  planar-quadrilateral pattern of single-panel meshes is not restated; such
  meshes take the generic colouring.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import Mesh2D, W, S, E, N

SPACES = ("W0", "W1", "W2", "W2H", "W2V", "W2broken", "W3", "WTHETA")


@dataclass(frozen=True)
class DofCounts:
    vert: int = 0
    edge_h: int = 0
    edge_v: int = 0
    face_h: int = 0
    face_v: int = 0
    vol: int = 0
    cell: int = 0


def ndof_counts(space: str, k_h: int = 0, k_v: int | None = None) -> DofCounts:
    """Dofs per entity (ndof_setup, function_space_constructor_helper_functions_mod.F90:225-306)."""
    if k_v is None:
        k_v = k_h
    h, v = k_h, k_v
    if space == "W0":
        return DofCounts(vert=1, edge_h=h, edge_v=v, face_h=h * v, face_v=h * h,
                         vol=h * h * v, cell=(h + 2) * (h + 2) * (v + 2))
    if space == "W1":
        return DofCounts(edge_h=h + 1, edge_v=v + 1,
                         face_h=(h + 1) * v + h * (v + 1), face_v=2 * (h + 1) * h,
                         vol=2 * h * (h + 1) * v + h * h * (v + 1),
                         cell=2 * (h + 1) * (h + 2) * (v + 2) + (h + 2) * (h + 2) * (v + 1))
    if space == "W2":
        return DofCounts(face_h=(h + 1) * (v + 1), face_v=(h + 1) * (h + 1),
                         vol=2 * h * (h + 1) * (v + 1) + (h + 1) * (h + 1) * v,
                         cell=2 * (h + 2) * (h + 1) * (v + 1) + (h + 1) * (h + 1) * (v + 2))
    if space == "W2H":
        return DofCounts(face_h=(h + 1) * (v + 1), vol=2 * h * (h + 1) * (v + 1),
                         cell=2 * (h + 1) * (h + 2) * (v + 1))
    if space == "W2V":
        return DofCounts(face_v=(h + 1) * (h + 1), vol=(h + 1) * (h + 1) * v,
                         cell=(h + 1) * (h + 1) * (v + 2))
    if space == "W2broken":
        nv = 2 * (h + 1) * (h + 2) * (v + 1) + (h + 1) * (h + 1) * (v + 2)
        return DofCounts(vol=nv, cell=nv)
    if space == "W3":
        nv = (h + 1) * (h + 1) * (v + 1)
        return DofCounts(vol=nv, cell=nv)
    if space == "WTHETA":
        return DofCounts(face_v=(h + 1) * (h + 1), vol=(h + 1) * (h + 1) * v,
                         cell=(h + 1) * (h + 1) * (v + 2))
    raise ValueError(f"unsupported space {space}")


def _selected_faces(space: str):
    """setup_select_entities: which faces of the cell carry dofs for the space."""
    if space in ("WTHETA", "W2V"):
        return (False,) * 4 + (True, True)
    if space == "W2H":
        return (True,) * 4 + (False, False)
    return (True,) * 6


def _side_face_keys(mesh: Mesh2D) -> np.ndarray:
    """Key per (cell, side) equal for the two cells sharing a side face."""
    nc = mesh.ncells
    cn = mesh.cell_next.astype(np.int64)              # 1-based, 0 = boundary
    own = np.arange(nc, dtype=np.int64)[:, None] * 4 + np.arange(4)[None, :]
    keys = own.copy()
    has = cn > 0
    nb = np.where(has, cn - 1, 0)
    # matching side on the neighbour: the slot whose neighbour is this cell
    match = cn[nb] == (np.arange(nc, dtype=np.int64)[:, None, None] + 1)
    nmatch = match.sum(axis=2)
    if np.any(nmatch[has] != 1):
        raise ValueError("mesh has a neighbour pair sharing more than one side")
    side_nb = match.argmax(axis=2)
    other = nb * 4 + side_nb
    keys[has] = np.minimum(own[has], other[has])
    return keys


@dataclass
class FunctionSpace:
    """A function space on an extruded serial mesh (the parts the path uses)."""

    space: str
    k: int
    nlayers: int
    mesh: Mesh2D
    counts: DofCounts
    dofmap: np.ndarray        # (ncells, ndf) int32, 1-based bottom-cell dofs
    undf: int
    dof_shift: np.ndarray     # (ndf,) 1 for dofs on a top entity (map = column start + 1)
    dof_height: np.ndarray    # (ndf,) dofs in the column the local dof walks (nl or nl+1)

    @property
    def ndf(self) -> int:
        return self.counts.cell

    @property
    def ncells(self) -> int:
        return self.mesh.ncells

    @property
    def ncell_3d(self) -> int:
        return self.mesh.ncells * self.nlayers

    @property
    def last_dof_owned(self) -> int:
        return self.undf

    @property
    def is_discontinuous(self) -> bool:
        c = self.counts
        return c.vert == c.edge_h == c.edge_v == c.face_h == c.face_v == 0

    def dofmap_fortran(self) -> np.ndarray:
        """dofmap(ndf, ncells) in Fortran (column-major) memory order."""
        return np.ascontiguousarray(self.dofmap)  # row c holds map(:, c)


def boundary_dofs(fs: "FunctionSpace") -> np.ndarray:
    """boundary_value(ndf, 2) of get_boundary_dofs: 0 for a dof on the bottom
    (column 1) / top (column 2) entity of the cell, 1 otherwise
    (function_space_constructor_helper_functions_mod.F90:678-680, 884-885,
    930-932: dofs of the bottom/top face, edges and vertices).  Those are
    exactly the dofs of nl+1-level columns: shift 0 bottom, shift 1 top."""
    bv = np.ones((fs.ndf, 2), dtype=np.int32)
    stack = fs.dof_height == fs.nlayers + 1
    bv[stack & (fs.dof_shift == 0), 0] = 0
    bv[stack & (fs.dof_shift == 1), 1] = 0
    return bv


def build_dofmap(mesh: Mesh2D, space: str, nlayers: int, k: int = 0,
                 k_v: int | None = None) -> FunctionSpace:
    """Serial dofmap_setup (all cells owned, ndata = 1)."""
    cnt = ndof_counts(space, k, k_v)
    nl = int(nlayers)
    nc = mesh.ncells
    sel = _selected_faces(space)
    fkeys = _side_face_keys(mesh)                                 # (nc, 4)
    vkeys = mesh.vert_on_cell.astype(np.int64)                    # (nc, 4)
    cells = np.arange(nc, dtype=np.int64)

    # Event slots per cell in traversal order: 0 vol | 1-4 side faces |
    # 5 bottom/top face pair | 6-9 horizontal edges | 10-13 vertical edges |
    # 14-17 vertices.  Entity keys live in disjoint ranges per entity kind.
    nslot = 18
    key = np.empty((nc, nslot), dtype=np.int64)
    size = np.zeros((nc, nslot), dtype=np.int64)
    big = np.int64(4 * nc + 8)
    key[:, 0] = 0 * big + cells                                   # volume
    size[:, 0] = cnt.vol * nl
    for f in range(4):
        key[:, 1 + f] = 1 * big + fkeys[:, f]
        size[:, 1 + f] = cnt.face_h * nl if sel[f] else 0
    key[:, 5] = 2 * big + cells
    size[:, 5] = cnt.face_v * (nl + 1) if (sel[4] or sel[5]) else 0
    for e in range(4):
        key[:, 6 + e] = 3 * big + fkeys[:, e]
        size[:, 6 + e] = cnt.edge_h * (nl + 1)
    for v in range(4):
        key[:, 10 + v] = 4 * big + vkeys[:, v]
        size[:, 10 + v] = cnt.edge_v * nl
        key[:, 14 + v] = 5 * big + vkeys[:, v]
        size[:, 14 + v] = cnt.vert * (nl + 1)

    flat_key = key.reshape(-1)
    flat_size = size.reshape(-1)
    uniq, first, inv = np.unique(flat_key, return_index=True, return_inverse=True)
    is_first = np.zeros(flat_key.size, dtype=bool)
    is_first[first] = True
    consumed = np.where(is_first, flat_size, 0)
    start = np.cumsum(consumed) - consumed + 1                    # exclusive prefix, 1-based
    base = start[first][inv].reshape(nc, nslot)                   # base of the entity's block
    undf = int(consumed.sum())

    cols, shift, height = [], [], []

    def emit(b, nd, stride, top):
        for i in range(nd):
            cols.append(b + i * stride + (1 if top else 0))
            shift.append(1 if top else 0)
            height.append(stride)

    emit(base[:, 0], cnt.vol, nl, False)
    for f in range(4):
        if sel[f]:
            emit(base[:, 1 + f], cnt.face_h, nl, False)
    if sel[4]:
        emit(base[:, 5], cnt.face_v, nl + 1, False)
    if sel[5]:
        emit(base[:, 5], cnt.face_v, nl + 1, True)
    for e in range(4):
        emit(base[:, 6 + e], cnt.edge_h, nl + 1, False)
    for e in range(4):
        emit(base[:, 10 + e], cnt.edge_v, nl, False)
    for e in range(4):
        emit(base[:, 6 + e], cnt.edge_h, nl + 1, True)
    for v in range(4):
        emit(base[:, 14 + v], cnt.vert, nl + 1, False)
    for v in range(4):
        emit(base[:, 14 + v], cnt.vert, nl + 1, True)

    dofmap = np.stack(cols, axis=1).astype(np.int32) if cols else np.zeros((nc, 0), np.int32)
    ndf = dofmap.shape[1]
    if space in ("W0", "W1", "W2", "W3", "W2broken") and ndf != cnt.cell:
        raise AssertionError(f"{space}: built {ndf} dofs per cell, expected {cnt.cell}")
    counts = cnt if ndf == cnt.cell else DofCounts(**{**cnt.__dict__, "cell": ndf})
    return FunctionSpace(space=space, k=k, nlayers=nl, mesh=mesh, counts=counts,
                         dofmap=dofmap, undf=undf,
                         dof_shift=np.array(shift, dtype=np.int32),
                         dof_height=np.array(height, dtype=np.int32))


# ---------------------------------------------------------------------------
# Colouring
# ---------------------------------------------------------------------------

@dataclass
class Colouring:
    ncolours: int
    ncells_per_colour: np.ndarray     # (ncolours,) int32
    cells_in_colour: np.ndarray       # (ncolours, maxcells) int32, 1-based, 0-padded
    colour_map: np.ndarray            # (ncells,) int32 colour of each cell (1-based)

    def order(self) -> np.ndarray:
        """1-based cells in the order the coloured PSy loop visits them."""
        return np.concatenate([self.cells_in_colour[c, :self.ncells_per_colour[c]]
                               for c in range(self.ncolours)])


# 2x2 colour patterns per panel, indexed [panel][column-1][row-1]
# (mesh_colouring_mod.F90:229-269; Fortran reshape fills column-major).
def _pattern(vals):
    return np.array(vals, dtype=np.int32).reshape(2, 2, order="F")


_EVEN = [_pattern(v) for v in ((1, 2, 3, 4), (3, 2, 1, 4), (3, 4, 1, 2),
                               (1, 4, 3, 2), (3, 2, 4, 1), (2, 3, 1, 4))]
_ODD = [_pattern(v) for v in ((1, 2, 3, 4), (4, 1, 2, 3), (1, 2, 3, 4),
                              (4, 1, 2, 3), (3, 2, 4, 1), (3, 2, 4, 1))]


def colours_ok(mesh: Mesh2D, colour_map: np.ndarray) -> bool:
    """test_set_colours (mesh_colouring_mod.F90:765-837): no edge or corner
    neighbour shares a cell's colour."""
    cm = np.concatenate([[0], colour_map])          # colour_map(0) = 0
    cn = mesh.cell_next
    col = cm[1:]
    nbr_cols = cm[cn]                                # (nc, 4)
    if np.any(nbr_cols == col[:, None]):
        return False
    ids = np.arange(1, mesh.ncells + 1)
    for d, back in ((N, S), (S, N)):
        t = cn[:, d]
        ok = t > 0
        tt = np.where(ok, t, 1) - 1
        straight = cn[tt, back] == ids
        lr = np.where(straight[:, None], cn[tt][:, [E, W]], cn[tt][:, [N, S]])
        clash = (cm[lr] == col[:, None]).any(axis=1) & ok
        if np.any(clash):
            return False
    return True


def _lookup(colour_map: np.ndarray, ncolours: int) -> Colouring:
    counts = np.array([(colour_map == c).sum() for c in range(1, ncolours + 1)], dtype=np.int32)
    cic = np.zeros((ncolours, int(counts.max())), dtype=np.int32)
    for c in range(1, ncolours + 1):
        cells = np.nonzero(colour_map == c)[0] + 1
        cic[c - 1, :cells.size] = cells
    return Colouring(ncolours=ncolours, ncells_per_colour=counts,
                     cells_in_colour=cic, colour_map=colour_map.astype(np.int32))


def _greedy(mesh: Mesh2D) -> np.ndarray:
    """set_colours_generic (mesh_colouring_mod.F90:514-650): cells in order take
    the smallest colour not used by their N, E, W, S neighbours and the two
    cells beside the N and S neighbours (orientation checked by going back)."""
    nc = mesh.ncells
    cn = mesh.cell_next.astype(np.int64)
    cmap = np.zeros(nc + 1, dtype=np.int32)            # cmap[0] = 0 for "no cell"
    for c in range(1, nc + 1):
        used = set()
        for d, back in ((N, S), (S, N)):
            t = cn[c - 1, d]
            if t > 0:
                used.add(int(cmap[t]))
                if cn[t - 1, back] == c:
                    used.add(int(cmap[cn[t - 1, E]])); used.add(int(cmap[cn[t - 1, W]]))
                else:
                    used.add(int(cmap[cn[t - 1, N]])); used.add(int(cmap[cn[t - 1, S]]))
        for d in (E, W):
            t = cn[c - 1, d]
            if t > 0:
                used.add(int(cmap[t]))
        col = 1
        while col in used:
            col += 1
        cmap[c] = col
    return cmap[1:]


def colour_cells(mesh: Mesh2D) -> Colouring:
    if mesh.npanels == 6:
        n = mesh.edge_cells
        cpp = n * n
        gid = mesh.global_cell_id.astype(np.int64)
        panel = (gid - 1) // cpp + 1
        rel = gid - cpp * (panel - 1)
        column = 2 - np.mod(np.mod(rel, n), 2)
        row = np.mod((rel - 1) // n, 2) + 1
        pats = _EVEN if n % 2 == 0 else _ODD
        table = np.stack(pats)                       # (6, 2, 2)
        cmap = table[panel - 1, column - 1, row - 1]
        if colours_ok(mesh, cmap):
            return _lookup(cmap, 4)
    # Odd-sized cubed spheres fail the 2x2 pattern's own check (the last and
    # second-to-last columns of a panel get the same colour) and, as in the
    # reference, fall back to the generic greedy colouring.  Single-panel
    # (planar) meshes use the generic colouring here too.
    cmap = _greedy(mesh)
    return _lookup(cmap, int(cmap.max()))
