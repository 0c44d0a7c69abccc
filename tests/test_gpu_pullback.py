"""GPU parity of the pullbacks (spec2grid_pullback!, grid2spec_pullback!,
SpeedyTransforms/ext/SpeedyTransformsEnzymeExt.jl:263-319) against the FP64 oracle's
restatement, which tests/test_oracle.py pins by the dot-product identity; the engine's own
outputs satisfy the identity with its forward transforms to FP32 rounding."""
import numpy as np
import pytest
import torch

from sht_cases import CONFIGS, rel
from test_gpu_transform import TOL, case, plan, to_dev

pytestmark = pytest.mark.gpu


def grid_cotangent(sw, S, K, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((K, S.grid.npoints))


@pytest.mark.parametrize("name", ["T31", "T127", "T31full"])
@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
@pytest.mark.parametrize("unscale", [False, True])
def test_spec2grid_pullback_vs_oracle(sw, gpu, name, precision, unscale):
    o, a, layers, fields = case(sw, name, seed=11)
    K = layers * fields
    S = plan(sw, name, precision)
    print(f"spec2grid_pullback {name} {precision} unscale={unscale}", flush=True)
    gbar = grid_cotangent(sw, S, K, 3)
    ref = o.spec2grid_pullback(gbar, unscale_coslat=unscale)
    fbar = sw.Field(torch.from_numpy(gbar.astype(np.float32)).cuda(), S.grid)
    cbar = sw.LowerTriangularArray(torch.zeros((K, S.spectrum.nharm), dtype=torch.complex64, device="cuda"),
                                   S.spectrum)
    sw.spec2grid_pullback_(cbar, fbar, S, unscale_coslat=unscale)
    torch.cuda.synchronize()
    got = cbar.data.cpu().numpy().astype(np.complex128)
    assert rel(got, ref) < TOL[precision]
    # accumulate (the reference's add=true) adds onto what is there; accumulate=False overwrites
    sw.spec2grid_pullback_(cbar, fbar, S, unscale_coslat=unscale)
    torch.cuda.synchronize()
    assert rel(cbar.data.cpu().numpy().astype(np.complex128), 2 * ref) < TOL[precision]
    sw.spec2grid_pullback_(cbar, fbar, S, unscale_coslat=unscale, accumulate=False)
    torch.cuda.synchronize()
    assert rel(cbar.data.cpu().numpy().astype(np.complex128), ref) < TOL[precision]


@pytest.mark.parametrize("name", ["T31", "T127", "T255", "T31full"])
@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
@pytest.mark.parametrize("kcols", [None, 5])
def test_grid2spec_pullback_vs_oracle(sw, gpu, name, precision, kcols):
    """kcols=5: a partial last column group in the accumulating FFT synthesis."""
    o, a, layers, fields = case(sw, name, seed=12)
    if kcols:
        a = a[:kcols]
    K = a.shape[0]
    S = plan(sw, name, precision)
    print(f"grid2spec_pullback {name} {precision} K={K}", flush=True)
    ref = o.grid2spec_pullback(a)
    cbar = to_dev(sw, a, S.spectrum)
    base = grid_cotangent(sw, S, K, 4) * np.abs(ref).max()  # of the result's size
    fbar = sw.Field(torch.from_numpy(base.astype(np.float32)).cuda(), S.grid)
    sw.grid2spec_pullback_(fbar, cbar, S)  # accumulates onto base
    torch.cuda.synchronize()
    assert rel(fbar.data.cpu().numpy().astype(np.float64) - base, ref) < 10 * TOL[precision]
    sw.grid2spec_pullback_(fbar, cbar, S, accumulate=False)
    torch.cuda.synchronize()
    assert rel(fbar.data.cpu().numpy().astype(np.float64), ref) < TOL[precision]


@pytest.mark.parametrize("name", ["T127", "T255"])
def test_engine_dot_product_identity(sw, gpu, name):
    """<g, S a> = Re<S^T g, a> and Re<b, A f> = <A^T b, f> with every term from the engine."""
    o, a, layers, fields = case(sw, name, seed=13)
    K = layers * fields
    S = plan(sw, name, "fp32")
    print(f"dot-product identity {name}", flush=True)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((K, S.grid.npoints))
    f = rng.standard_normal((K, S.grid.npoints))
    b = sw.seeded_spectrum(S.spectrum, K, seed=14)
    coeffs = to_dev(sw, a, S.spectrum)
    Sa = sw.transform(coeffs, S).data.double().cpu().numpy()
    gd = sw.Field(torch.from_numpy(g.astype(np.float32)).cuda(), S.grid)
    STg = sw.spec2grid_pullback_(sw.LowerTriangularArray(torch.zeros_like(coeffs.data), S.spectrum), gd, S,
                                 accumulate=False).data.cpu().numpy().astype(np.complex128)
    fd = sw.Field(torch.from_numpy(f.astype(np.float32)).cuda(), S.grid)
    Af = sw.transform(fd, S).data.cpu().numpy().astype(np.complex128)
    bd = to_dev(sw, b, S.spectrum)
    ATb = sw.grid2spec_pullback_(sw.Field(torch.zeros_like(fd.data), S.grid), bd, S,
                                 accumulate=False).data.double().cpu().numpy()
    torch.cuda.synchronize()
    g32, a32 = g.astype(np.float32).astype(np.float64), a.astype(np.complex64).astype(np.complex128)
    lhs, rhs = np.sum(g32 * Sa), np.real(np.sum(np.conj(STg) * a32))
    assert abs(lhs - rhs) < 1e-5 * np.sqrt(np.sum(g32 ** 2) * np.sum(Sa ** 2))
    f32, b32 = f.astype(np.float32).astype(np.float64), b.astype(np.complex64).astype(np.complex128)
    lhs, rhs = np.real(np.sum(np.conj(b32) * Af)), np.sum(ATb * f32)
    assert abs(lhs - rhs) < 1e-5 * np.sqrt(np.sum(np.abs(b32) ** 2) * np.sum(np.abs(Af) ** 2))
