"""Ring grids and fields (RingGrids/src/grids/octahedral_gaussian.jl:23-116,
full_gaussian.jl:17-61, field.jl:7-18) and the resolution rules of
SpeedyTransforms/src/aliasing.jl:9-69."""
from __future__ import annotations

import math

import numpy as np
import torch

from ._lib import DimensionMismatch


def roundup_fft(n: int, small_primes=(2, 3, 5)) -> int:
    """Smallest even m >= n with only small prime factors (aliasing.jl:42-59)."""
    n += n % 2
    while True:
        r = n
        for p in small_primes:
            while r % p == 0:
                r //= p
        if r == 1:
            return n
        n += 2


def get_nlat_half(truncation: int, dealiasing: float = 2.0) -> int:
    """aliasing.jl:9-15 (truncation 1-based)."""
    return roundup_fft(math.ceil(((1 + dealiasing) * (truncation - 1) + 1) / 4))


def get_truncation(nlat_half: int, dealiasing: float = 2.0) -> int:
    """aliasing.jl:23-29."""
    return math.floor((4 * nlat_half - 1) / (dealiasing + 1)) + 1


class _RingGrid:
    kind = ""

    def __init__(self, nlat_half: int):
        if nlat_half < 1:
            raise ValueError("nlat_half must be >= 1")
        self.nlat_half = int(nlat_half)

    def nlon_ring(self, j: int) -> int:  # 0-based northern ring
        raise NotImplementedError

    @property
    def nlat(self) -> int:
        return 2 * self.nlat_half

    @property
    def nlons(self) -> np.ndarray:
        north = [self.nlon_ring(j) for j in range(self.nlat_half)]
        return np.array(north + north[::-1])

    @property
    def npoints(self) -> int:
        return int(self.nlons.sum())

    @property
    def nlon_max(self) -> int:
        return int(self.nlons.max())

    def eachring(self):
        """0-based half-open ranges of every ring, north to south (eachring, grid.jl:173-184)."""
        starts = np.concatenate([[0], np.cumsum(self.nlons)])
        return [range(int(starts[i]), int(starts[i + 1])) for i in range(self.nlat)]

    def __eq__(self, other):
        return type(self) is type(other) and self.nlat_half == other.nlat_half

    def __repr__(self):
        return f"{type(self).__name__}({self.nlat_half})"


class OctahedralGaussianGrid(_RingGrid):
    """16 + 4j points on northern ring j = 1..nlat_half, mirrored south, no offset."""
    kind = "OctahedralGaussianGrid"

    def nlon_ring(self, j: int) -> int:
        return 16 + 4 * (j + 1)


class FullGaussianGrid(_RingGrid):
    """4 nlat_half points on every ring."""
    kind = "FullGaussianGrid"

    def nlon_ring(self, j: int) -> int:
        return 4 * self.nlat_half


class Field:
    """Gridded data. `data` is a torch tensor of shape (npoints,) or (K, npoints): the memory
    order of the reference's column-major `field.data[ij, k]`."""

    def __init__(self, data: torch.Tensor, grid: _RingGrid):
        if data.shape[-1] != grid.npoints:
            raise DimensionMismatch(f"data has {data.shape[-1]} points, {grid} has {grid.npoints}")
        self.data, self.grid = data, grid

    @classmethod
    def zeros(cls, grid: _RingGrid, K: int | None = None, dtype=torch.float32, device="cpu"):
        shape = (grid.npoints,) if K is None else (K, grid.npoints)
        return cls(torch.zeros(shape, dtype=dtype, device=device), grid)

    @property
    def nlayers(self) -> int:
        return 1 if self.data.dim() == 1 else self.data.shape[0]

    def __repr__(self):
        return f"Field({self.grid}, K={self.nlayers}, {self.data.dtype}, {self.data.device})"
