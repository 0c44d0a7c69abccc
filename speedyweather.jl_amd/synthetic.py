"""Seeded synthetic spectra: SplitMix64 -> 53-bit uniforms -> Box-Muller.

Column k of seed s uses the SplitMix64 stream started at state s * 1000003 + k.  Draws
follow the packed (m-major) order, two uniforms (real, imaginary) per coefficient;
coefficients are scaled by (l+1)^-1.5 and the imaginary part is zeroed at m = 0."""
from __future__ import annotations

import numpy as np

from .spectrum import Spectrum

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(state: int, count: int) -> np.ndarray:
    """`count` outputs of SplitMix64 from `state` (uint64)."""
    with np.errstate(over="ignore"):
        s = np.uint64(state) + _GOLDEN * np.arange(1, count + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniforms(state: int, count: int) -> np.ndarray:
    return (splitmix64(state, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def seeded_spectrum(spectrum: Spectrum, K: int, seed: int, decay: float = 1.5) -> np.ndarray:
    """(K, nharm) complex128 coefficients, Re, Im ~ N(0,1) (l+1)^-decay, Im = 0 at m = 0."""
    ls, ms = spectrum.degrees_orders()
    out = np.empty((K, spectrum.nharm), np.complex128)
    amp = (ls + 1.0) ** (-decay)
    for k in range(K):
        u = uniforms((seed * 1000003 + k) % (1 << 64), 2 * spectrum.nharm).reshape(-1, 2)
        r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
        th = 2.0 * np.pi * u[:, 1]
        z = r * np.cos(th) + 1j * (r * np.sin(th))
        z[ms == 0] = z[ms == 0].real
        out[k] = z * amp
    return out


def seeded_grid(npoints: int, K: int, seed: int) -> np.ndarray:
    """(K, npoints) float64 N(0,1) grid values from the same generator."""
    out = np.empty((K, npoints))
    for k in range(K):
        u = uniforms((seed * 1000003 + k + (1 << 40)) % (1 << 64), 2 * ((npoints + 1) // 2)).reshape(-1, 2)
        r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
        out[k] = (r * np.cos(2.0 * np.pi * u[:, 1]))[:npoints]
    return out
