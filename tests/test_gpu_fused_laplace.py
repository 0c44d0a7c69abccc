"""GPU parity of the transforms fused with the spectral Laplacian (sht_synthesis_laplace,
sht_analysis_laplace): transform(∇² a) and ∇² transform(f) (spectral_gradients.jl:385-466
composed with transform!, spectral_transform.jl:515-631) against the FP64 oracle's
transform and Laplacian, and against the unfused GPU composition of the two calls."""
import numpy as np
import pytest
import torch

from oracle.sht_oracle import laplace as o_laplace
from sht_cases import CONFIGS, oracle, rel
from test_gpu_transform import plan, to_dev

pytestmark = pytest.mark.gpu

OPTS = [(False, False, False, 1.0), (True, False, False, 1.0), (False, True, True, 2.5), (True, True, True, 6.371e6)]


def host(x):
    return x.data.cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("name", ["T31", "T127"])
@pytest.mark.parametrize("inverse, flipsign, add, radius", OPTS)
def test_fused_laplace_vs_oracle(sw, gpu, name, inverse, flipsign, add, radius):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    K = 8
    S = plan(sw, name, "fp32", nlayers=K)
    o = oracle(kind, n, lmax, mmax)
    print(f"fused laplace {name} inverse={inverse} flip={flipsign} add={add} R={radius}", flush=True)
    a = sw.seeded_spectrum(S.spectrum, K, seed=31)
    lap = o_laplace(o.spectrum, a, inverse=inverse, flipsign=flipsign, radius=radius)
    # spectral -> grid of the Laplacian (add has no meaning there)
    f = sw.Field(torch.zeros((K, S.grid.npoints), device="cuda"), S.grid)
    sw.transform_laplace_(f, to_dev(sw, a, S.spectrum), S, inverse=inverse, flipsign=flipsign, radius=radius)
    fu = sw.Field(torch.zeros((K, S.grid.npoints), device="cuda"), S.grid)
    sw.transform_laplace_(fu, to_dev(sw, a, S.spectrum), S, inverse=inverse, flipsign=flipsign, radius=radius,
                          unscale_coslat=True)
    # grid -> spectral, then the Laplacian (with add into o0)
    g = o.synthesis(sw.seeded_spectrum(S.spectrum, K, seed=32))
    o0 = sw.seeded_spectrum(S.spectrum, K, seed=33)
    ref_b = o_laplace(o.spectrum, o.analysis(g), inverse=inverse, flipsign=flipsign, add=o0 if add else None,
                      radius=radius)
    b = to_dev(sw, o0, S.spectrum)
    sw.transform_laplace_(b, sw.Field(torch.from_numpy(g.astype(np.float32)).cuda(), S.grid), S, inverse=inverse,
                          flipsign=flipsign, add=add, radius=radius)
    torch.cuda.synchronize()
    assert rel(f.data.cpu().numpy(), o.synthesis(lap)) < 1e-5
    assert rel(fu.data.cpu().numpy(), o.synthesis(lap, unscale_coslat=True)) < 1e-5
    assert rel(host(b), ref_b) < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
@pytest.mark.parametrize("inverse, flipsign, add, radius", OPTS[::3])
def test_fused_laplace_matches_unfused_t255(sw, gpu, precision, inverse, flipsign, add, radius):
    """At T255 K=64 the fused calls agree with transform + laplace_ run as separate calls."""
    S = plan(sw, "T255", precision)
    K = S.nlayers
    print(f"fused vs unfused T255 {precision} inverse={inverse} add={add}", flush=True)
    a = sw.seeded_spectrum(S.spectrum, K, seed=41)
    c = to_dev(sw, a, S.spectrum)
    lap = sw.laplace_(sw.LowerTriangularArray(torch.zeros_like(c.data), S.spectrum), c, S, inverse=inverse,
                      flipsign=flipsign, radius=radius)
    f_sep = sw.transform(lap, S)
    f_fused = sw.Field(torch.zeros_like(f_sep.data), S.grid)
    sw.transform_laplace_(f_fused, c, S, inverse=inverse, flipsign=flipsign, radius=radius)
    o0 = to_dev(sw, sw.seeded_spectrum(S.spectrum, K, seed=42), S.spectrum)
    b_sep = o0.data.clone()
    spec = sw.transform(f_sep, S)
    sw.laplace_(sw.LowerTriangularArray(b_sep, S.spectrum), spec, S, inverse=inverse, flipsign=flipsign, add=add,
                radius=radius)
    b_fused = o0.data.clone()
    sw.transform_laplace_(sw.LowerTriangularArray(b_fused, S.spectrum), f_sep, S, inverse=inverse, flipsign=flipsign,
                          add=add, radius=radius)
    torch.cuda.synchronize()
    assert rel(f_fused.data.cpu().numpy(), f_sep.data.cpu().numpy()) < 1e-5
    assert rel(b_fused.cpu().numpy(), b_sep.cpu().numpy()) < 1e-5


def test_fused_laplace_on_matrix_plan_and_errors(sw, gpu):
    """A dense plan runs the fused calls on the transform kernels; bad arguments raise."""
    kind, n, lmax, mmax, _, _ = CONFIGS["T31"]
    o = oracle(kind, n, lmax, mmax)
    M = sw.MatrixSpectralTransform(sw.Spectrum(lmax, mmax), sw.OctahedralGaussianGrid(n), nlayers=4)
    a = sw.seeded_spectrum(M.spectrum, 4, seed=51)
    f = sw.Field(torch.zeros((4, M.grid.npoints), device="cuda"), M.grid)
    sw.transform_laplace_(f, to_dev(sw, a, M.spectrum), M)
    torch.cuda.synchronize()
    assert rel(f.data.cpu().numpy(), o.synthesis(o_laplace(o.spectrum, a))) < 1e-5
    with pytest.raises(sw.SHTError):
        sw.transform_laplace_(f, to_dev(sw, a, M.spectrum), M, radius=0.0)
    with pytest.raises(TypeError):
        sw.transform_laplace_(f, to_dev(sw, a, M.spectrum), M, add=True)
    M.close()
