"""Legendre shortcuts on the GPU (SpeedyTransforms/src/legendre_shortcuts.jl:21-49) against
the FP64 oracle: all five, including the latitude-dependent ones that are not monotonic in
the ring index on full Gaussian grids (a ring nearer the equator keeps fewer orders than
one nearer the pole; the reference's own GPU test of that case is
SpeedyWeather/test/GPU/spectral_transform.jl:297-322)."""
import numpy as np
import pytest
import torch

from sht_cases import BANDS, rel
from test_gpu_transform import TOL, to_dev
from oracle.sht_oracle import Grid, OracleTransform, Spectrum as OSpectrum

pytestmark = pytest.mark.gpu

CASES = [
    ("full", 24, 33, 32, "lincubcoslat", "fp32"),
    ("full", 24, 33, 32, "lincubcoslat", "fp16x3"),
    ("full", 24, 33, 32, "lincubcoslat", "bf16x3"),
    ("full", 48, 65, 64, "lincubcoslat", "fp32"),
    ("octahedral", 24, 33, 32, "quadratic", "fp32"),
    ("octahedral", 24, 33, 32, "cubic", "fp16x3"),
    ("octahedral", 96, 129, 128, "lincubcoslat", "fp32"),
    ("octahedral", 192, 257, 256, "linquadcoslat2", "fp16x3"),
]


@pytest.mark.parametrize("kind, n, lmax, mmax, shortcut, precision", CASES)
def test_shortcut_transforms_vs_oracle(sw, gpu, kind, n, lmax, mmax, shortcut, precision):
    print(f"shortcut case {kind} n={n} {shortcut} {precision}", flush=True)
    o = OracleTransform(OSpectrum(lmax, mmax), Grid(kind, n), shortcut)
    if kind == "full" and shortcut == "lincubcoslat":
        assert (np.diff(o.mtrunc) < 0).any()
    K = 6
    grid = (sw.OctahedralGaussianGrid if kind == "octahedral" else sw.FullGaussianGrid)(n)
    sp = sw.Spectrum(lmax, mmax)
    S = sw.SpectralTransform(sp, grid, nlayers=K, LegendreShortcut=shortcut, precision=precision)
    a = sw.seeded_spectrum(sp, K, seed=21)
    coeffs = to_dev(sw, a, sp)
    f = sw.transform(coeffs, S).data.cpu().numpy()
    f_ref = o.synthesis(a)
    assert rel(f, f_ref) < TOL[precision]
    # analysis of a grid with content at every order a ring holds (not only the kept ones)
    g = np.random.default_rng(22).standard_normal((K, grid.npoints))
    b = sw.transform(sw.Field(torch.from_numpy(g.astype(np.float32)).cuda(), grid), S).data.cpu().numpy()
    b_ref = o.analysis(g.astype(np.float32).astype(np.float64))
    for lo, hi in BANDS:
        sel = (o.spectrum.l_of >= lo) & (o.spectrum.l_of <= hi)
        if sel.any():
            assert rel(b[:, sel], b_ref[:, sel]) < TOL[precision], (lo, hi)
    # the pullback of the synthesis goes through the same FFT analysis rows
    gb = sw.Field(torch.from_numpy(g.astype(np.float32)).cuda(), grid)
    cb = sw.LowerTriangularArray(torch.zeros((K, sp.nharm), dtype=torch.complex64, device="cuda"), sp)
    sw.spec2grid_pullback_(cb, gb, S, accumulate=False)
    assert rel(cb.data.cpu().numpy(), o.spec2grid_pullback(g.astype(np.float32).astype(np.float64))) < TOL[precision]
    S.close()


def test_shortcut_plans_in_batches_and_graph_replay(sw, gpu):
    """A batch that mixes a plan with a non-monotonic shortcut (full grid, LinCubCoslat) and a
    linear-shortcut plan, and graph replay of single calls on the shortcut plan, equal direct
    single calls bitwise (the zero Fourier rows up to the running maximum included)."""
    K = 5
    sp = sw.Spectrum(33, 32)
    Sf = sw.SpectralTransform(sp, sw.FullGaussianGrid(24), nlayers=K, LegendreShortcut="lincubcoslat")
    So = sw.SpectralTransform(sp, sw.OctahedralGaussianGrid(24), nlayers=K)
    items, expect = [], []
    for i, S in enumerate((Sf, So)):
        a = sw.seeded_spectrum(sp, K, seed=40 + i)
        c = sw.LowerTriangularArray(torch.from_numpy(a).to(torch.complex64).cuda(), sp)
        mid = sw.Field(torch.zeros((K, S.grid.npoints), device="cuda"), S.grid)
        out = sw.LowerTriangularArray.zeros(sp, K, device="cuda")
        items.append((S, c, out, mid))
        f = sw.transform(c, S)
        expect.append((f.data.clone(), sw.transform(f, S).data.clone()))
    B = sw.SpectralBatch(items)
    B.launch()
    torch.cuda.synchronize()
    for (S, c, out, mid), (f, b) in zip(items, expect):
        assert torch.equal(mid.data, f) and torch.equal(out.data, b)
    B.close()
    Sf.set_graphs(True)
    c = items[0][1]
    f = sw.Field(torch.zeros((K, Sf.grid.npoints), device="cuda"), Sf.grid)
    b = sw.LowerTriangularArray.zeros(sp, K, device="cuda")
    for rep in range(3):
        sw.transform_(f, c, Sf)
        sw.transform_(b, f, Sf)
        torch.cuda.synchronize()
        assert torch.equal(f.data, expect[0][0]) and torch.equal(b.data, expect[0][1]), rep
    assert Sf.graph_count() == (2, 0)
    Sf.close()
    So.close()
