"""Synthetic problems of the benchmark configurations (host arrays, numpy).

All random inputs come from ``rng.Stream`` (SplitMix64) so the CPU oracle
and the GPU see identical bytes.  Draw order is part of each definition:

* ``spd_lma(stream, n3, ndf, shift, std)``: for every cell-layer ik in order,
  a ndf x ndf matrix B (row-major draws, N(0, std^2)); the local matrix is
  B B^T + shift*I.  Stored as local_stencil(ncell_3d, ndf, ndf): numpy shape
  (ndf, ndf, ncell_3d) (identical memory to the Fortran array).
* ``normal_lma(stream, n3, ndf_to, ndf_from)``: entries N(0,1) drawn for
  ik-major, then row (to), then column (from).

Configuration "precision" (C96, 70 layers, quadratic extrusion to 80 km):
W2 mass = BB^T + 6I, B ~ N(0,1), seed 1; then div (W2 -> W3, 1x6) and grad
(W3 -> W2, 6x1) ~ N(0,1); then the right-hand side b ~ U[-1,1] on W2.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import function_space as fsm
from .mesh import cubed_sphere, extrusion_eta
from .rng import Stream


def spd_lma(stream: Stream, n3: int, ndf: int, shift: float, std: float = 1.0,
            dtype=np.float64, chunk: int = 1 << 18) -> np.ndarray:
    out = np.empty((ndf, ndf, n3), dtype=dtype)
    eye = shift * np.eye(ndf)
    for s in range(0, n3, chunk):
        m = min(chunk, n3 - s)
        B = stream.normal(m * ndf * ndf).reshape(m, ndf, ndf) * std
        A = np.einsum("kab,kcb->kac", B, B) + eye          # A[k, i, j]
        out[:, :, s:s + m] = A.transpose(2, 1, 0)             # out[j, i, k] = A[k, i, j]
    return out


def normal_lma(stream: Stream, n3: int, ndf_to: int, ndf_from: int, dtype=np.float64,
               chunk: int = 1 << 18) -> np.ndarray:
    out = np.empty((ndf_from, ndf_to, n3), dtype=dtype)
    for s in range(0, n3, chunk):
        m = min(chunk, n3 - s)
        A = stream.normal(m * ndf_to * ndf_from).reshape(m, ndf_to, ndf_from)
        out[:, :, s:s + m] = A.transpose(2, 1, 0)
    return out


@dataclass
class PrecisionConfig:
    """W2/W3 lowest-order operators on a cubed sphere (the C96L70 precision case)."""

    edge_cells: int
    nlayers: int
    mesh: object
    w2: fsm.FunctionSpace
    w3: fsm.FunctionSpace
    colouring: fsm.Colouring
    eta: np.ndarray
    mass: np.ndarray      # (6, 6, n3) float64
    div: np.ndarray       # (6, 1, n3): local_stencil(n3, 1, 6)
    grad: np.ndarray      # (1, 6, n3): local_stencil(n3, 6, 1)
    b: np.ndarray         # (undf_w2,) float64


def precision_config(edge_cells: int = 96, nlayers: int = 70, seed: int = 1,
                     with_div_grad: bool = True) -> PrecisionConfig:
    mesh = cubed_sphere(edge_cells)
    w2 = fsm.build_dofmap(mesh, "W2", nlayers)
    w3 = fsm.build_dofmap(mesh, "W3", nlayers)
    col = fsm.colour_cells(mesh)
    n3 = mesh.ncells * nlayers
    st = Stream(seed)
    mass = spd_lma(st, n3, 6, 6.0)
    if with_div_grad:
        div = normal_lma(st, n3, 1, 6)
        grad = normal_lma(st, n3, 6, 1)
    else:
        div = grad = None
    b = st.uniform(w2.undf)
    return PrecisionConfig(edge_cells, nlayers, mesh, w2, w3, col,
                           80.0e3 * extrusion_eta(nlayers, "quadratic"), mass, div, grad, b)


@dataclass
class DiffusionConfig:
    """Assembled W0 operator on a cubed-sphere shell (the C12L10 plain-method case)."""

    edge_cells: int
    nlayers: int
    mesh: object
    w0: fsm.FunctionSpace
    colouring: fsm.Colouring
    heights: np.ndarray
    matrix: np.ndarray    # (8, 8, n3) float64: mass + dt_kappa * stiffness
    mass: np.ndarray
    stiffness: np.ndarray
    x_true: np.ndarray    # (undf_w0,) float64


def diffusion_config(node_lon_deg, node_lat_deg, face_nodes, nlayers: int = 10,
                     depth: float = 10.0e3, radius: float = 6.371e6, dt_kappa: float = 1.0e-3,
                     seed: int = 0, n_q: int = 3) -> DiffusionConfig:
    """W0 lowest order (8 dofs/cell) on the shell given by a cubed-sphere mesh file's
    node longitudes/latitudes (degrees) and face-node table (1-based), ``nlayers``
    uniform layers over ``depth`` metres; local matrices mass + dt_kappa * stiffness
    by ``n_q``-point Gaussian quadrature; x_true ~ U[-1, 1] from seed ``seed``."""
    from .assembly import shell_corners, to_local_stencil, w0_mass_stiffness
    ncell = face_nodes.shape[0]
    edge = int(round(np.sqrt(ncell / 6)))
    mesh = cubed_sphere(edge)
    w0 = fsm.build_dofmap(mesh, "W0", nlayers)
    col = fsm.colour_cells(mesh)
    heights = depth * extrusion_eta(nlayers, "uniform")
    chi = shell_corners(node_lon_deg, node_lat_deg, face_nodes, radius, heights)
    mass, stiff = w0_mass_stiffness(chi, n_q)
    A = mass + dt_kappa * stiff
    x_true = Stream(seed).uniform(w0.undf)
    return DiffusionConfig(edge, nlayers, mesh, w0, col, heights, to_local_stencil(A),
                           to_local_stencil(mass), to_local_stencil(stiff), x_true)


def spd_lma_device(stream, n3: int, ndf: int, shift: float, std: float, chunk: int = 1 << 16):
    """spd_lma drawn from a rng.TorchStream on its device (float64 tensor
    (ndf, ndf, n3)); same draw order and construction as ``spd_lma``."""
    import torch
    out = torch.empty((ndf, ndf, n3), dtype=torch.float64, device=stream.device)
    eye = shift * torch.eye(ndf, dtype=torch.float64, device=stream.device)
    for s in range(0, n3, chunk):
        m = min(chunk, n3 - s)
        B = stream.normal(m * ndf * ndf).reshape(m, ndf, ndf) * std
        A = torch.bmm(B, B.transpose(1, 2).contiguous()) + eye
        out[:, :, s:s + m] = A.permute(2, 1, 0)
    return out


@dataclass
class MultiConfig:
    """Element-order-1 W2 mass operator and 16 fields (the C48L70 multi-field case)."""

    edge_cells: int
    nlayers: int
    mesh: object
    w2: fsm.FunctionSpace
    colouring: fsm.Colouring
    matrix: object        # float64 (36, 36, n3) on the device
    x: object             # float64 (undf, 16) on the device


def multi_config(device, edge_cells: int = 48, nlayers: int = 70, seed: int = 3) -> MultiConfig:
    """W2 k=1 (36 dofs/cell) local matrices B B^T + 36 I, B ~ N(0, 1/36); then
    16 fields ~ U[-1, 1] drawn dof-major ((undf, 16) row-major), one stream."""
    from .rng import TorchStream
    mesh = cubed_sphere(edge_cells)
    w2 = fsm.build_dofmap(mesh, "W2", nlayers, k=1)
    col = fsm.colour_cells(mesh)
    st = TorchStream(seed, device)
    n3 = mesh.ncells * nlayers
    M = spd_lma_device(st, n3, w2.ndf, float(w2.ndf), 1.0 / np.sqrt(w2.ndf))
    x = st.uniform(w2.undf * 16).reshape(w2.undf, 16)
    return MultiConfig(edge_cells, nlayers, mesh, w2, col, M, x)
