"""Writes the golden fixtures of the parity tests with the FP64 oracle.

T31 (lmax=33, mmax=32) on the octahedral Gaussian grid nlat_half=24, 8 columns, seeds 0
and 1 of the package's seeded (l+1)^-1.5 spectrum: input coefficients, the FP64
spectral->grid output, and the FP64 grid->spectral output of that grid (the round trip).
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)

import __graft_entry__  # noqa: E402
from oracle.sht_oracle import Grid, OracleTransform, Spectrum  # noqa: E402


def main():
    sw = __graft_entry__.load_package()
    o = OracleTransform(Spectrum(33, 32), Grid("octahedral", 24))
    arrays = {}
    for seed in (0, 1):
        a = sw.seeded_spectrum(sw.Spectrum(33, 32), 8, seed=seed)
        f = o.synthesis(a)
        arrays[f"coeffs_{seed}"] = a
        arrays[f"grid_{seed}"] = f
        arrays[f"roundtrip_{seed}"] = o.analysis(f)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "t31_octahedral_k8.npz")
    np.savez_compressed(path, **arrays)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
