"""Golden fixtures (tests/golden/t31_octahedral_k8.npz, written by make_golden.py):
the oracle and the generator reproduce them on the CPU; the engine matches them on the GPU."""
import os

import numpy as np
import pytest

from sht_cases import oracle, rel

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "t31_octahedral_k8.npz")


@pytest.fixture(scope="module")
def fix():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_generator_reproduces_inputs(sw, fix):
    for seed in (0, 1):
        a = sw.seeded_spectrum(sw.Spectrum(33, 32), 8, seed=seed)
        assert np.array_equal(a, fix[f"coeffs_{seed}"])


def test_oracle_reproduces_fixture(fix):
    o = oracle("octahedral", 24, 33, 32)
    for seed in (0, 1):
        f = o.synthesis(fix[f"coeffs_{seed}"])
        assert np.abs(f - fix[f"grid_{seed}"]).max() < 1e-13
        b = o.analysis(f)
        assert np.abs(b - fix[f"roundtrip_{seed}"]).max() < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp16x3", "bf16x3"])
def test_engine_matches_fixture(sw, gpu, fix, precision):
    import torch
    sp, grid = sw.Spectrum(33, 32), sw.OctahedralGaussianGrid(24)
    S = sw.SpectralTransform(sp, grid, nlayers=8, precision=precision)
    tol = 2e-6
    for seed in (0, 1):
        c = sw.LowerTriangularArray(torch.from_numpy(fix[f"coeffs_{seed}"]).to(torch.complex64).cuda(), sp)
        f = sw.transform(c, S)
        b = sw.transform(f, S)
        torch.cuda.synchronize()
        assert rel(f.data.cpu().numpy(), fix[f"grid_{seed}"]) < tol
        assert rel(b.data.cpu().numpy(), fix[f"roundtrip_{seed}"]) < tol
