"""GPU parity of the spectral gradient operators (divergence!, curl!, ∇²!, ∇⁻²!, ∇!,
SpeedyTransforms/src/spectral_gradients.jl:30-125,385-527) against the FP64 oracle's
restatement, which tests/test_oracle.py pins by finite differences of the fields."""
import numpy as np
import pytest
import torch

from oracle.sht_oracle import divergence as o_divergence, gradient as o_gradient, laplace as o_laplace
from sht_cases import CONFIGS, oracle, rel
from test_gpu_transform import plan, to_dev

pytestmark = pytest.mark.gpu


def arrays(sw, S, K, seeds):
    return [sw.seeded_spectrum(S.spectrum, K, seed=s) for s in seeds]


def zeros(sw, S, K):
    return sw.LowerTriangularArray(torch.zeros((K, S.spectrum.nharm), dtype=torch.complex64, device="cuda"),
                                   S.spectrum)


def host(x):
    return x.data.cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("name", ["T31", "T127", "T255"])
@pytest.mark.parametrize("curl", [False, True])
@pytest.mark.parametrize("flipsign, add, radius", [(False, False, 1.0), (True, True, 6.371e6)])
def test_divergence_curl_vs_oracle(sw, gpu, name, curl, flipsign, add, radius):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    S = plan(sw, name, "fp32")
    K = layers * fields
    print(f"{'curl' if curl else 'divergence'} {name} flip={flipsign} add={add} R={radius}", flush=True)
    u, v, o0 = arrays(sw, S, K, (1, 2, 3))
    ref = o_divergence(oracle(kind, n, lmax, mmax).spectrum, u, v, curl=curl, flipsign=flipsign,
                       add=o0 if add else None, radius=radius)
    out = to_dev(sw, o0, S.spectrum)
    (sw.curl_ if curl else sw.divergence_)(out, to_dev(sw, u, S.spectrum), to_dev(sw, v, S.spectrum), S,
                                          flipsign=flipsign, add=add, radius=radius)
    torch.cuda.synchronize()
    assert rel(host(out), ref) < 1e-6


@pytest.mark.parametrize("name", ["T31", "T127", "T255"])
@pytest.mark.parametrize("inverse, flipsign, add, radius", [(False, False, False, 1.0), (True, False, False, 1.0),
                                                            (False, True, True, 2.5), (True, True, True, 6.371e6)])
def test_laplace_vs_oracle(sw, gpu, name, inverse, flipsign, add, radius):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    S = plan(sw, name, "fp32")
    K = layers * fields
    print(f"laplace {name} inverse={inverse} flip={flipsign} add={add} R={radius}", flush=True)
    a, o0 = arrays(sw, S, K, (4, 5))
    ref = o_laplace(oracle(kind, n, lmax, mmax).spectrum, a, inverse=inverse, flipsign=flipsign,
                    add=o0 if add else None, radius=radius)
    out = to_dev(sw, o0, S.spectrum)
    sw.laplace_(out, to_dev(sw, a, S.spectrum), S, add=add, flipsign=flipsign, inverse=inverse, radius=radius)
    torch.cuda.synchronize()
    assert rel(host(out), ref) < 1e-6
    # in place
    x = to_dev(sw, a, S.spectrum)
    sw.laplace_(x, x, S, inverse=inverse)
    torch.cuda.synchronize()
    assert rel(host(x), o_laplace(oracle(kind, n, lmax, mmax).spectrum, a, inverse=inverse)) < 1e-6


@pytest.mark.parametrize("name", ["T31", "T127", "T255"])
@pytest.mark.parametrize("radius", [1.0, 6.371e6])
def test_gradient_vs_oracle(sw, gpu, name, radius):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    S = plan(sw, name, "fp32")
    K = layers * fields
    print(f"gradient {name} R={radius}", flush=True)
    (p,) = arrays(sw, S, K, (6,))
    rx, ry = o_gradient(oracle(kind, n, lmax, mmax).spectrum, p, radius=radius)
    dx, dy = zeros(sw, S, K), zeros(sw, S, K)
    sw.gradient_(dx, dy, to_dev(sw, p, S.spectrum), S, radius=radius)
    torch.cuda.synchronize()
    assert rel(host(dx), rx) < 1e-6 and rel(host(dy), ry) < 1e-6


def test_operator_argument_errors(sw, gpu):
    S = plan(sw, "T31", "fp32")
    u, v = (to_dev(sw, x, S.spectrum) for x in arrays(sw, S, 8, (1, 2)))
    with pytest.raises(sw.SHTError):
        sw.divergence_(u, u, v, S)  # output aliases an input
    with pytest.raises(sw.SHTError):
        sw.gradient_(u, v, u, S)
    with pytest.raises(sw.DimensionMismatch):
        sw.divergence_(zeros(sw, S, 4), u, v, S)
