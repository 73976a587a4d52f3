"""LFRic UGRID mesh files: read the 2D mesh topologies and intergrid maps.

The reference reads its meshes through ``ncdf_quad_type``
(infrastructure/source/io/ncdf_quad_mod.F90): ``scan_for_topologies``
(:3465) lists the variables whose ``cf_role`` is ``mesh_topology``,
``get_dimensions`` (:1897-1958) and ``read_mesh`` (:2023-2539) take, for
one named mesh, its attributes (geometry, topology, coord_sys, periodic_x/y,
n_mesh_maps, maps_to), the node and face coordinates
``<mesh>_node_x/_node_y/_face_x/_face_y``, the connectivities
``<mesh>_face_nodes``, ``<mesh>_face_edges``, ``<mesh>_face_links`` (with
its ``_FillValue`` as the void cell) and ``<mesh>_edge_nodes``, and
``read_map`` (:2557-2683) the cell map ``<source>_<target>_map`` of
dimensions ``n<source>_map_face`` x ``n<target>_cells_per_<source>_y`` x
``n<target>_cells_per_<source>_x``.  This module reads the same variables
from the file's NetCDF-3 encoding (classic, 64-bit offset or CDF-5), parsed
here directly (the format is the published NetCDF "classic" specification;
no netCDF library is needed), and hands the topology to
``mesh.from_ugrid_arrays`` so the result feeds ``function_space`` exactly as
the built-in generator's meshes do.

Arrays keep the file's (C, row-major) order: a Fortran ``face_nodes(4,
nfaces)`` is ``(nfaces, 4)`` here, ``cell_map(x, y, cell)`` is
``(ncells, y, x)``.  Connectivities are returned 1-based whatever the file's
``start_index``.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .mesh import Mesh2D, from_ugrid_arrays

_ABSENT = 0
_NC_DIMENSION = 0x0A
_NC_VARIABLE = 0x0B
_NC_ATTRIBUTE = 0x0C
# nc_type -> numpy big-endian dtype (NC_CHAR handled as text)
_TYPES = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8",
          7: ">u1", 8: ">u2", 9: ">u4", 10: ">i8", 11: ">u8"}


class UgridError(ValueError):
    """A file that is not a NetCDF-3 file or not an LFRic UGRID mesh file."""


@dataclass
class NcVariable:
    name: str
    dims: tuple
    attrs: dict
    data: np.ndarray


@dataclass
class NcFile:
    dims: dict
    attrs: dict
    variables: dict


class _Reader:
    def __init__(self, buf: bytes):
        self.buf = buf
        self.pos = 0
        magic = buf[:4]
        if len(buf) < 8 or magic[:3] != b"CDF" or magic[3] not in (1, 2, 5):
            raise UgridError("not a NetCDF-3 file (magic %r)" % magic)
        self.version = magic[3]
        self.pos = 4
        self.wide = self.version == 5            # 8-byte counts and dimension ids
        self.offset8 = self.version in (2, 5)    # 8-byte variable offsets

    def _take(self, n: int) -> bytes:
        if self.pos + n > len(self.buf):
            raise UgridError("truncated NetCDF header")
        b = self.buf[self.pos:self.pos + n]
        self.pos += n
        return b

    def u32(self) -> int:
        return struct.unpack(">I", self._take(4))[0]

    def u64(self) -> int:
        return struct.unpack(">Q", self._take(8))[0]

    def count(self) -> int:
        return self.u64() if self.wide else self.u32()

    def name(self) -> str:
        n = self.count()
        s = self._take(n)
        self._take((4 - n % 4) % 4)
        return s.decode("utf-8")

    def values(self, nc_type: int, n: int):
        if nc_type not in _TYPES:
            raise UgridError("unknown nc_type %d" % nc_type)
        dt = np.dtype(_TYPES[nc_type])
        nbytes = n * dt.itemsize
        raw = self._take(nbytes)
        self._take((4 - nbytes % 4) % 4)
        if nc_type == 2:
            return raw.rstrip(b"\x00").decode("utf-8", errors="replace")
        v = np.frombuffer(raw, dtype=dt).astype(dt.newbyteorder("="))
        return v[0].item() if n == 1 else v

    def list_header(self, tag: int):
        t = self.u32()
        n = self.count()
        if t == _ABSENT:
            if n != 0:
                raise UgridError("malformed absent list")
            return 0
        if t != tag:
            raise UgridError("expected list tag %d, found %d" % (tag, t))
        return n

    def attrs(self) -> dict:
        out = {}
        for _ in range(self.list_header(_NC_ATTRIBUTE)):
            nm = self.name()
            nc_type = self.u32()
            n = self.count()
            out[nm] = self.values(nc_type, n)
        return out


def read_netcdf(path) -> NcFile:
    """Every dimension, global attribute and variable of a NetCDF-3 file."""
    buf = Path(path).read_bytes()
    r = _Reader(buf)
    numrecs = r.count()
    dims = []
    for _ in range(r.list_header(_NC_DIMENSION)):
        nm = r.name()
        dims.append((nm, r.count()))
    gattrs = r.attrs()
    vars_ = []
    for _ in range(r.list_header(_NC_VARIABLE)):
        nm = r.name()
        nd = r.count()
        ids = [r.count() for _ in range(nd)]
        va = r.attrs()
        nc_type = r.u32()
        vsize = r.count()
        begin = r.u64() if r.offset8 else r.u32()
        vars_.append((nm, ids, va, nc_type, vsize, begin))
    for _, ids, _, _, _, _ in vars_:
        if any(i >= len(dims) for i in ids):
            raise UgridError("variable refers to a missing dimension")
    # record variables: those whose first dimension has length 0 (unlimited)
    is_rec = [bool(ids) and dims[ids[0]][1] == 0 for _, ids, _, _, _, _ in vars_]
    recvars = [v for v, rec in zip(vars_, is_rec) if rec]
    recsize = sum(v[4] for v in recvars)
    if len(recvars) == 1:
        # a lone record variable is not padded between records
        v = recvars[0]
        dt = np.dtype(_TYPES[v[3]])
        recsize = int(np.prod([dims[i][1] for i in v[1][1:]], dtype=np.int64)) * dt.itemsize
    variables = {}
    for (nm, ids, va, nc_type, vsize, begin), rec in zip(vars_, is_rec):
        dt = np.dtype(_TYPES[nc_type])
        shape = [dims[i][1] for i in ids]
        if rec:
            shape[0] = numrecs
            inner = int(np.prod(shape[1:], dtype=np.int64)) * dt.itemsize
            chunks = []
            for k in range(numrecs):
                a = begin + k * recsize
                if a + inner > len(buf):
                    raise UgridError("record of %s runs past the end of the file" % nm)
                chunks.append(buf[a:a + inner])
            raw = b"".join(chunks)
        else:
            nbytes = int(np.prod(shape, dtype=np.int64)) * dt.itemsize
            if begin + nbytes > len(buf):
                raise UgridError("variable %s runs past the end of the file" % nm)
            raw = buf[begin:begin + nbytes]
        if nc_type == 2:
            data = np.frombuffer(raw, dtype="S1").reshape(shape)
        else:
            data = np.frombuffer(raw, dtype=dt).astype(dt.newbyteorder("=")).reshape(shape)
        variables[nm] = NcVariable(nm, tuple(dims[i][0] for i in ids), va, data)
    return NcFile(dict(dims), gattrs, variables)


def _text(v) -> str:
    return v if isinstance(v, str) else str(v)


def mesh_names(nc: NcFile) -> list:
    """Mesh topologies in file order (``scan_for_topologies``: cf_role =
    mesh_topology)."""
    return [n for n, v in nc.variables.items() if _text(v.attrs.get("cf_role", "")).strip() == "mesh_topology"]


def _flag(v) -> bool:
    # Fortran logicals written as text ('T', 'F', '.true.', ...)
    s = _text(v).strip().strip(".").upper()
    return s.startswith("T")


@dataclass
class UgridMesh:
    """One mesh of an LFRic UGRID file, as ``read_mesh`` returns it."""

    name: str
    geometry: str
    topology: str
    coord_sys: str
    periodic_xy: tuple
    coord_units: tuple
    node_coordinates: np.ndarray        # (2, nnodes) x/lon, y/lat
    face_coordinates: np.ndarray        # (2, nfaces)
    face_node_connectivity: np.ndarray  # (nfaces, 4), 1-based
    face_edge_connectivity: np.ndarray  # (nfaces, 4), 1-based
    face_face_connectivity: np.ndarray  # (nfaces, 4), 1-based, void_cell where no neighbour
    edge_node_connectivity: np.ndarray  # (nedges, 2), 1-based
    void_cell: int
    maps_to: list = field(default_factory=list)
    attrs: dict = field(default_factory=dict)

    @property
    def nfaces(self) -> int:
        return int(self.face_node_connectivity.shape[0])

    @property
    def npanels(self) -> int:
        """6 for a cubed sphere (spherical, 6 n^2 faces), otherwise 1."""
        n = int(round(np.sqrt(self.nfaces / 6)))
        return 6 if self.geometry == "spherical" and 6 * n * n == self.nfaces else 1

    def to_mesh2d(self) -> Mesh2D:
        """The serial local mesh (cell_next in W,S,E,N order, corners in
        SW,SE,NE,NW order) that function_space builds dofmaps from."""
        links = np.where(self.face_face_connectivity == self.void_cell, 0, self.face_face_connectivity)
        return from_ugrid_arrays(links, self.face_node_connectivity, self.npanels, name=self.name)


def _conn(v: NcVariable) -> np.ndarray:
    a = np.asarray(v.data, dtype=np.int64)
    start = v.attrs.get("start_index", 1)
    start = int(start if not isinstance(start, np.ndarray) else start[0])
    fill = v.attrs.get("_FillValue")
    if start == 0:
        a = np.where(a == fill, a, a + 1) if fill is not None else a + 1
    return a.astype(np.int32)


def read_mesh(path, name: str | None = None) -> UgridMesh:
    """The mesh `name` (default: the first topology) of a UGRID file."""
    nc = read_netcdf(path)
    names = mesh_names(nc)
    if not names:
        raise UgridError("%s holds no mesh topology" % path)
    if name is None:
        name = names[0]
    if name not in names:
        raise KeyError("mesh %r not in %s (meshes: %s)" % (name, path, ", ".join(names)))
    top = nc.variables[name].attrs

    def var(suffix):
        key = name + suffix
        if key not in nc.variables:
            raise UgridError("mesh %r has no variable %s" % (name, key))
        return nc.variables[key]

    def coords(nm):
        ref = _text(top.get(nm, "")).split()
        if len(ref) == 2 and all(r in nc.variables for r in ref):
            return [nc.variables[r] for r in ref]
        return [var("_node_x"), var("_node_y")] if nm == "node_coordinates" else [var("_face_x"), var("_face_y")]

    nx, ny = coords("node_coordinates")
    fx, fy = coords("face_coordinates")
    links = var("_face_links")
    fill = links.attrs.get("_FillValue", 0)
    void = int(fill if not isinstance(fill, np.ndarray) else fill[0])
    n_maps = int(top.get("n_mesh_maps", 0) or 0)
    maps_to = _text(top.get("maps_to", "")).split()[:n_maps] if n_maps else []
    face_face = _conn(links)
    return UgridMesh(
        name=name,
        geometry=_text(top.get("geometry", "")).strip(),
        topology=_text(top.get("topology", "")).strip(),
        coord_sys=_text(top.get("coord_sys", "")).strip(),
        periodic_xy=(_flag(top.get("periodic_x", "F")), _flag(top.get("periodic_y", "F"))),
        coord_units=(_text(nx.attrs.get("units", "")), _text(ny.attrs.get("units", ""))),
        node_coordinates=np.stack([np.asarray(nx.data, np.float64), np.asarray(ny.data, np.float64)]),
        face_coordinates=np.stack([np.asarray(fx.data, np.float64), np.asarray(fy.data, np.float64)]),
        face_node_connectivity=_conn(var("_face_nodes")),
        face_edge_connectivity=_conn(var("_face_edges")),
        face_face_connectivity=face_face,
        edge_node_connectivity=_conn(var("_edge_nodes")),
        void_cell=void,
        maps_to=maps_to,
        attrs=dict(top),
    )


def read_map(path, source: str, target: str) -> np.ndarray:
    """cell_map of `source` onto `target`: (ncells_source, ny, nx), the
    1-based target cells covering each source cell (``read_map``)."""
    nc = read_netcdf(path)
    key = "%s_%s_map" % (source, target)
    if key not in nc.variables:
        raise KeyError("no map %s in %s" % (key, path))
    v = nc.variables[key]
    want = ("n%s_map_face" % source, "n%s_cells_per_%s_y" % (target, source),
            "n%s_cells_per_%s_x" % (target, source))
    if v.dims != want:
        raise UgridError("map %s has dimensions %s, expected %s" % (key, v.dims, want))
    return np.asarray(v.data, dtype=np.int32)


def load_mesh(path, name: str | None = None) -> Mesh2D:
    """Shorthand: the serial Mesh2D of mesh `name` in a UGRID file."""
    return read_mesh(path, name).to_mesh2d()
