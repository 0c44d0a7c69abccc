"""Spectral layout: Spectrum and LowerTriangularArray
(LowerTriangularArrays/src/spectrum.jl:6-70, lower_triangular_array.jl:10-20,367-395)."""
from __future__ import annotations

import math

import numpy as np
import torch

from ._lib import DimensionMismatch


class Spectrum:
    """1-based spectral truncation: lmax degrees and mmax orders (spectrum.jl:37-70)."""

    def __init__(self, lmax: int, mmax: int | None = None):
        mmax = lmax if mmax is None else mmax
        if not (1 <= mmax <= lmax):
            raise ValueError("need 1 <= mmax <= lmax")
        self.lmax, self.mmax = int(lmax), int(mmax)

    @classmethod
    def from_truncation(cls, truncation: int, one_more_degree: bool = False) -> "Spectrum":
        """Spectrum(truncation; one_degree_more) (spectrum.jl:72-78)."""
        return cls(truncation + int(one_more_degree), truncation)

    @property
    def nharm(self) -> int:
        """nonzeros (spectrum.jl:87-89)."""
        return self.lmax * self.mmax - self.mmax * (self.mmax - 1) // 2

    def offset(self, m: int) -> int:
        """0-based packed start of order m (0-based)."""
        return m * self.lmax - m * (m - 1) // 2

    def lm2i(self, l: int, m: int) -> int:
        """0-based packed index of 0-based degree l and order m (lm2i, :367)."""
        if not (0 <= m < self.mmax and m <= l < self.lmax):
            raise IndexError((l, m))
        return self.offset(m) + l - m

    def degrees_orders(self):
        ls, ms = [], []
        for m in range(self.mmax):
            ls.extend(range(m, self.lmax))
            ms.extend([m] * (self.lmax - m))
        return np.array(ls), np.array(ms)

    def __eq__(self, other):
        return isinstance(other, Spectrum) and (self.lmax, self.mmax) == (other.lmax, other.mmax)

    def __repr__(self):
        return f"Spectrum(lmax={self.lmax}, mmax={self.mmax})"


class LowerTriangularArray:
    """Packed spherical harmonic coefficients. `data` is a torch tensor of shape (nharm,)
    or (K, nharm): the memory order of the reference's column-major `data[lm, k]`."""

    def __init__(self, data: torch.Tensor, spectrum: Spectrum):
        if data.shape[-1] != spectrum.nharm:
            raise DimensionMismatch(f"data has {data.shape[-1]} coefficients, {spectrum} needs {spectrum.nharm}")
        self.data, self.spectrum = data, spectrum

    @classmethod
    def zeros(cls, spectrum: Spectrum, K: int | None = None, dtype=torch.complex64, device="cpu"):
        shape = (spectrum.nharm,) if K is None else (K, spectrum.nharm)
        return cls(torch.zeros(shape, dtype=dtype, device=device), spectrum)

    @property
    def nlayers(self) -> int:
        return 1 if self.data.dim() == 1 else self.data.shape[0]

    def __repr__(self):
        return f"LowerTriangularArray({self.spectrum}, K={self.nlayers}, {self.data.dtype}, {self.data.device})"
