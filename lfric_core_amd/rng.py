"""Seeded, compiler-independent input generator.

The reference fills random fields with the Fortran intrinsic
``random_number`` (infrastructure/source/utilities/random_number_generator_mod.f90),
whose stream depends on the compiler and its seeding.  Inputs here come from
SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): a 64-bit Weyl sequence with increment
0x9E3779B97F4A7C15 followed by the published two-multiply finaliser.  The
same bytes are handed to the CPU oracle and to the GPU, so both always see
identical inputs.

* ``uniform01``: top 53 bits / 2**53, in [0, 1).
* ``uniform(a, b)``: a + (b - a) * uniform01.
* ``normal``: Box-Muller (Box & Muller 1958) on consecutive pairs
  (u1, u2) -> sqrt(-2 ln(1 - u1)) * cos(2 pi u2) and ... * sin(2 pi u2).
"""

from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Return outputs ``offset .. offset+n-1`` of SplitMix64 started at ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential consumer of one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.pos = 0

    def raw(self, n: int) -> np.ndarray:
        out = splitmix64(self.seed, n, self.pos)
        self.pos += n
        return out

    def uniform01(self, n: int) -> np.ndarray:
        return (self.raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def uniform(self, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
        return lo + (hi - lo) * self.uniform01(n)

    def normal(self, n: int, mean: float = 0.0, std: float = 1.0) -> np.ndarray:
        m = (n + 1) // 2
        u = self.uniform01(2 * m).reshape(m, 2)
        rad = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
        ang = 2.0 * np.pi * u[:, 1]
        z = np.empty((m, 2))
        z[:, 0] = rad * np.cos(ang)
        z[:, 1] = rad * np.sin(ang)
        return mean + std * z.reshape(-1)[:n]


_I64 = lambda u: u - (1 << 64) if u >= (1 << 63) else u


class TorchStream:
    """The same SplitMix64 stream and transforms as ``Stream``, evaluated on a
    torch device (int64 arithmetic wraps like uint64; logical shifts are
    masked arithmetic shifts).  Raw draws are bitwise those of ``Stream``;
    uniforms are exact; normals may differ from numpy's in the last bit of
    log/cos/sin."""

    def __init__(self, seed: int, device, chunk: int = 1 << 26):
        self.seed = int(seed)
        self.pos = 0
        self.device = device
        self.chunk = chunk

    @staticmethod
    def _srl(z, s: int):
        import torch
        return torch.bitwise_and(z >> s, (1 << (64 - s)) - 1)

    def raw(self, n: int):
        import torch
        out = torch.empty(n, dtype=torch.int64, device=self.device)
        g, m1, m2 = _I64(0x9E3779B97F4A7C15), _I64(0xBF58476D1CE4E5B9), _I64(0x94D049BB133111EB)
        seed = _I64(self.seed & 0xFFFFFFFFFFFFFFFF)
        for s in range(0, n, self.chunk):
            m = min(self.chunk, n - s)
            z = torch.arange(self.pos + s + 1, self.pos + s + m + 1, dtype=torch.int64,
                             device=self.device)
            z = z * g + seed
            z = (z ^ self._srl(z, 30)) * m1
            z = (z ^ self._srl(z, 27)) * m2
            out[s:s + m] = z ^ self._srl(z, 31)
        self.pos += n
        return out

    def uniform01(self, n: int):
        import torch
        return self._srl(self.raw(n), 11).to(torch.float64) * (1.0 / 9007199254740992.0)

    def uniform(self, n: int, lo: float = -1.0, hi: float = 1.0):
        return lo + (hi - lo) * self.uniform01(n)

    def normal(self, n: int, mean: float = 0.0, std: float = 1.0):
        import torch
        m = (n + 1) // 2
        u = self.uniform01(2 * m).reshape(m, 2)
        rad = torch.sqrt(-2.0 * torch.log1p(-u[:, 0]))
        ang = 2.0 * np.pi * u[:, 1]
        z = torch.stack([rad * torch.cos(ang), rad * torch.sin(ang)], dim=1)
        return mean + std * z.reshape(-1)[:n]
