"""The ring DFT on MFMA (a radix-4 pass and a p-point DFT as two GEMMs, FP32 products or
scaled FP16 hi/lo products) as the Fourier stage: transforms and pullbacks against the
FP64 oracle, with 4 and 8 columns per workgroup, partial column groups included."""
import numpy as np
import pytest
import torch

from sht_cases import CONFIGS, oracle, rel
from test_gpu_transform import case, to_dev

pytestmark = pytest.mark.gpu

KINDS = [("fp32", "dft32"), ("fp16x3", "dft32"), ("fp16x3", "dft16"), ("bf16x3", "dft16")]


def fresh_plan(sw, name, precision, K):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    grid = (sw.OctahedralGaussianGrid if kind == "octahedral" else sw.FullGaussianGrid)(n)
    return sw.SpectralTransform(sw.Spectrum(lmax, mmax), grid, nlayers=K, precision=precision)


@pytest.mark.parametrize("name", ["T31", "T127", "T255", "T31full"])
@pytest.mark.parametrize("precision, fourier", KINDS)
@pytest.mark.parametrize("cols", [4, 8])
def test_dft_fourier_stage_vs_oracle(sw, gpu, name, precision, fourier, cols):
    o, a, layers, fields = case(sw, name, seed=4)
    K = a.shape[0]
    print(f"dft case {name} {precision} {fourier} cols={cols} K={K}", flush=True)
    S = fresh_plan(sw, name, precision, K)
    S.set_tuning(fourier=fourier, dft_cols=cols)
    assert S.tuning()["fourier"] == fourier and S.tuning()["dft_cols"] == cols
    f_ref = o.synthesis(a)
    b_ref = o.analysis(f_ref)
    coeffs = to_dev(sw, a, S.spectrum)
    f = sw.transform(coeffs, S).data.cpu().numpy()
    fin = sw.Field(torch.from_numpy(f_ref.astype(np.float32)).cuda(), S.grid)
    b = sw.transform(fin, S).data.cpu().numpy()
    assert np.isfinite(f).all() and np.isfinite(b).all()
    assert rel(f, f_ref) < 1e-5, rel(f, f_ref)
    assert rel(b, b_ref) < 1e-5, rel(b, b_ref)
    # a partial column group (K not a multiple of the columns per workgroup) and unscale_coslat
    k5 = 5
    f5 = sw.transform(sw.LowerTriangularArray(coeffs.data[:k5], S.spectrum), S, unscale_coslat=True).data.cpu().numpy()
    assert rel(f5, o.synthesis(a[:k5], unscale_coslat=True)) < 1e-5
    b5 = sw.transform(sw.Field(fin.data[:k5], S.grid), S).data.cpu().numpy()
    assert rel(b5, b_ref[:k5]) < 1e-5
    S.close()


@pytest.mark.parametrize("name", ["T127", "T255"])
@pytest.mark.parametrize("precision, fourier", [("fp32", "dft32"), ("fp16x3", "dft16")])
def test_dft_pullbacks_vs_oracle(sw, gpu, name, precision, fourier):
    """The adjoint scalings and the accumulating synthesis of the DFT kernels."""
    o, a, layers, fields = case(sw, name, seed=5)
    K = 5
    a = a[:K]
    S = fresh_plan(sw, name, precision, K)
    S.set_tuning(fourier=fourier, dft_cols=4)
    rng = np.random.default_rng(6)
    gbar = rng.standard_normal((K, S.grid.npoints))
    pc = o.spec2grid_pullback(gbar)
    pf = o.grid2spec_pullback(a)
    # accumulate onto base values of the results' size
    base_c = (rng.standard_normal(pc.shape) + 1j * rng.standard_normal(pc.shape)) * np.abs(pc).max()
    base_f = rng.standard_normal(pf.shape) * np.abs(pf).max()
    cb = sw.LowerTriangularArray(torch.from_numpy(base_c).to(torch.complex64).cuda(), S.spectrum)
    sw.spec2grid_pullback_(cb, sw.Field(torch.from_numpy(gbar.astype(np.float32)).cuda(), S.grid), S)
    assert rel(cb.data.cpu().numpy(), base_c.astype(np.complex64) + pc) < 1e-5
    fb = sw.Field(torch.from_numpy(base_f.astype(np.float32)).cuda(), S.grid)
    sw.grid2spec_pullback_(fb, to_dev(sw, a, S.spectrum), S)
    assert rel(fb.data.cpu().numpy(), base_f.astype(np.float32) + pf) < 1e-5
    S.close()
