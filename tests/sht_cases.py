"""Shared cases and error measures for the parity tests."""
import functools

import numpy as np

from oracle.sht_oracle import Grid, OracleTransform, Spectrum as OSpectrum

# (name, grid kind, nlat_half, lmax, mmax, layers, fields)
CONFIGS = {
    "T31": ("octahedral", 24, 33, 32, 8, 1),
    "T127": ("octahedral", 96, 129, 128, 8, 4),
    "T255": ("octahedral", 192, 257, 256, 16, 4),
    "T31full": ("full", 24, 33, 32, 8, 1),
}
BANDS = [(0, 15), (16, 63), (64, 127), (128, 10 ** 9)]


@functools.lru_cache(maxsize=None)
def oracle(kind, n, lmax, mmax):
    return OracleTransform(OSpectrum(lmax, mmax), Grid(kind, n))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def per_field_band_errors(spec_out, spec_ref, spectrum_l, layers, fields):
    """{(field, band): relative L2 error} over (l, m, layers) of each field and degree band."""
    out = {}
    for f in range(fields):
        cols = slice(f * layers, (f + 1) * layers)
        for b, (lo, hi) in enumerate(BANDS):
            sel = (spectrum_l >= lo) & (spectrum_l <= hi)
            if not sel.any():
                continue
            out[(f, b)] = rel(spec_out[cols][:, sel], spec_ref[cols][:, sel])
    return out
