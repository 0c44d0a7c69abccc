"""GPU parity of the dense matrix transform (MatrixSpectralTransform,
SpeedyTransforms/src/matrix_transform.jl:56-260) against the FP64 oracle and the regular
transform: the matrices are the same linear maps, so both directions agree with the oracle
to FP32 rounding; the size guard refuses matrices above 4 GiB."""
import numpy as np
import pytest
import torch

from sht_cases import oracle, rel
from test_gpu_transform import to_dev

pytestmark = pytest.mark.gpu

CASES = {"T31": ("octahedral", 24, 33, 32), "T31full": ("full", 24, 33, 32), "T63": ("octahedral", 48, 65, 64)}


def make(sw, name, K, cls):
    kind, n, lmax, mmax = CASES[name]
    grid = (sw.OctahedralGaussianGrid if kind == "octahedral" else sw.FullGaussianGrid)(n)
    return cls(sw.Spectrum(lmax, mmax), grid, nlayers=K)


@pytest.mark.parametrize("name", ["T31", "T31full", "T63"])
def test_matrix_transform_vs_oracle(sw, gpu, name):
    kind, n, lmax, mmax = CASES[name]
    K = 8
    o = oracle(kind, n, lmax, mmax)
    M = make(sw, name, K, sw.MatrixSpectralTransform)
    S = make(sw, name, K, sw.SpectralTransform)
    print(f"matrix transform {name}", flush=True)
    a = sw.seeded_spectrum(M.spectrum, K, seed=21)
    f_ref = o.synthesis(a)
    b_ref = o.analysis(f_ref)
    c = to_dev(sw, a, M.spectrum)
    f = sw.transform(c, M)
    fin = sw.Field(torch.from_numpy(f_ref.astype(np.float32)).cuda(), M.grid)
    b = sw.transform(fin, M)
    fs = sw.transform(c, S)
    fu = sw.transform(c, M, unscale_coslat=True)
    torch.cuda.synchronize()
    assert rel(f.data.cpu().numpy(), f_ref) < 1e-5
    assert rel(b.data.cpu().numpy(), b_ref) < 1e-5
    assert rel(f.data.cpu().numpy(), fs.data.cpu().numpy()) < 1e-5
    assert rel(fu.data.cpu().numpy(), o.synthesis(a, unscale_coslat=True)) < 1e-5
    # K subsets and strided views take the same GEMMs with other sizes and strides
    wide = torch.full((3, M.grid.npoints + 7), 3.0, device="cuda")
    fv = sw.Field(wide[:, :M.grid.npoints], M.grid)
    sw.transform_(fv, sw.LowerTriangularArray(c.data[:3], M.spectrum), M)
    torch.cuda.synchronize()
    assert rel(fv.data.cpu().numpy(), f_ref[:3]) < 1e-5
    assert torch.all(wide[:, M.grid.npoints:] == 3.0)
    M.close()
    S.close()


@pytest.mark.parametrize("K", [20, 9, 1])
def test_matrix_transform_column_groups(sw, gpu, K):
    """Column counts past one group of the split-K GEMM (8 or 16 columns per workgroup),
    and a single column, in both directions, on a plan holding 20 layers."""
    kind, n, lmax, mmax = CASES["T31"]
    o = oracle(kind, n, lmax, mmax)
    M = make(sw, "T31", 20, sw.MatrixSpectralTransform)
    print(f"matrix transform T31 K={K}", flush=True)
    a = sw.seeded_spectrum(M.spectrum, K, seed=23)
    f_ref = o.synthesis(a)
    c = to_dev(sw, a, M.spectrum)
    f = sw.transform(c, M)
    fu = sw.transform(c, M, unscale_coslat=True)
    b = sw.transform(sw.Field(torch.from_numpy(f_ref.astype(np.float32)).cuda(), M.grid), M)
    torch.cuda.synchronize()
    assert rel(f.data.cpu().numpy(), f_ref) < 1e-5
    assert rel(fu.data.cpu().numpy(), o.synthesis(a, unscale_coslat=True)) < 1e-5
    assert rel(b.data.cpu().numpy(), o.analysis(f_ref)) < 1e-5
    M.close()


def test_matrix_transform_size_guard(sw, gpu):
    with pytest.raises(sw.SHTError):
        sw.MatrixSpectralTransform(sw.Spectrum(257, 256), sw.OctahedralGaussianGrid(192), nlayers=1)
