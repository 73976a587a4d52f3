"""Extract golden data from the reference tree into small fixtures.

Run in a container where the reference is mounted (default /root/reference):

    python tests/golden/make_fixtures.py [reference_root]

Writes, next to this script:

* ``fs_test_data.npz`` - the integer dofmaps and real nodal coordinates /
  basis tables of infrastructure/unit-test/support/fs_test_data_mod.F90
  (golden values the reference's own function-space tests assert against,
  3x3 biperiodic unit-test mesh, 3 layers).
* ``meshes.npz`` - face-face and face-node tables of the reference's
  cubed-sphere mesh files (UGRID, NetCDF-3), read with scipy's NetCDF reader,
  node and face coordinates of C12, and the intergrid cell maps of the
  multigrid mesh file (C32/C16/C8/C4, C12/C4).

Only data is written; no reference source is copied.
"""

from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

DECL = re.compile(
    r"(integer|real)\((?:kind=)?(\w+)\),\s*public,\s*target\s*::\s*(\w+)\(([^)]*)\)\s*=\s*reshape\(\s*\[",
    re.I)


def _dims(spec: str):
    out = []
    for d in spec.split(","):
        d = d.strip()
        if ":" in d:
            lo, hi = (int(v) for v in d.split(":"))
            out.append(hi - lo + 1)
        else:
            out.append(int(d))
    return out


def parse_fs_test_data(path: Path) -> dict:
    text = path.read_text()
    arrays = {}
    for m in DECL.finditer(text):
        kind, _, name, spec = m.groups()
        start = m.end()
        end = text.index("]", start)
        body = text[start:end].replace("&", " ")
        toks = [t for t in re.split(r"[,\s]+", body) if t]
        if kind.lower() == "integer":
            vals = np.array([int(t) for t in toks], dtype=np.int64)
        else:
            vals = np.array([float(re.sub(r"_r_def$", "", t).replace("D", "E").replace("d", "e"))
                             for t in toks], dtype=np.float64)
        shape = _dims(spec)
        if vals.size != int(np.prod(shape)):
            raise ValueError(f"{name}: {vals.size} values for shape {shape}")
        arrays[name] = vals.reshape(shape, order="F")
    return arrays


MESHES = {
    "C4": "infrastructure/unit-test/data/mesh_C4.nc",
    "C8": "infrastructure/unit-test/data/mesh_C8.nc",
    "C12": "applications/coupled/example/mesh_C12.nc",
    "C16": "infrastructure/unit-test/data/mesh_C16.nc",
    "C24": "applications/skeleton/example/mesh_C24.nc",
    "C32": "infrastructure/unit-test/data/mesh_C32.nc",
}


def read_meshes(root: Path) -> dict:
    from scipy.io import netcdf_file
    out = {}
    for tag, rel in MESHES.items():
        f = netcdf_file(str(root / rel), "r", mmap=False)
        names = list(f.variables)
        pick = lambda suffix: [k for k in names if k.endswith(suffix)][0]
        out[f"{tag}_face_links"] = f.variables[pick("face_links")].data.astype(np.int32).copy()
        out[f"{tag}_face_nodes"] = f.variables[pick("face_nodes")].data.astype(np.int32).copy()
        if tag == "C12":
            for v in ("node_x", "node_y", "face_x", "face_y"):
                out[f"{tag}_{v}"] = f.variables[pick(v)].data.astype(np.float64).copy()
        f.close()
    return out


MG_FILE = "infrastructure/unit-test/data/mesh_C32_MG.nc"


def read_mg_maps(root: Path) -> dict:
    """Intergrid cell maps of the reference's multigrid mesh file, stored as
    cell_map(x, y, source cell) (Fortran order) -> numpy [cell, y, x]."""
    from scipy.io import netcdf_file
    f = netcdf_file(str(root / MG_FILE), "r", mmap=False)
    out = {k: f.variables[k].data.astype(np.int32).copy() for k in f.variables if k.endswith("_map")}
    f.close()
    return out


def main(root: str = "/root/reference") -> None:
    r = Path(root)
    fs = parse_fs_test_data(r / "infrastructure/unit-test/support/fs_test_data_mod.F90")
    np.savez_compressed(HERE / "fs_test_data.npz", **fs)
    np.savez_compressed(HERE / "meshes.npz", **read_meshes(r), **read_mg_maps(r))
    print("wrote", sorted(fs)[:5], "...", len(fs), "arrays")


if __name__ == "__main__":
    main(*sys.argv[1:])
