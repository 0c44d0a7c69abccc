"""The autotuner's choices (FFT factorisation policies, launch orders) keep parity: every
pinned configuration reproduces the FP64 oracle's transforms, and autotune picks one of
them, caches it per size class and leaves results within FP32 rounding of the default."""
import itertools

import numpy as np
import pytest
import torch

from sht_cases import CONFIGS, oracle, rel
from test_gpu_transform import case, to_dev

pytestmark = pytest.mark.gpu

POLICIES = list(itertools.product([17, 11], [0, 1, 2], [False, True]))


def fresh_plan(sw, name, precision="fp32"):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    grid = (sw.OctahedralGaussianGrid if kind == "octahedral" else sw.FullGaussianGrid)(n)
    return sw.SpectralTransform(sw.Spectrum(lmax, mmax), grid, nlayers=layers * fields, precision=precision)


@pytest.mark.parametrize("name, precision", [("T127", "fp32"), ("T31full", "fp32"), ("T255", "fp32"),
                                             ("T255", "fp16x3")])
def test_every_pinned_configuration_matches_oracle(sw, gpu, name, precision):
    """At T255 this covers the MFMA prime pass with more row tiles than waves (primes of
    131 and above occur only on its rings) under every policy."""
    o, a, layers, fields = case(sw, name, seed=1)
    f_ref = o.synthesis(a)
    b_ref = o.analysis(f_ref)
    S = fresh_plan(sw, name, precision)
    coeffs = to_dev(sw, a, S.spectrum)
    fin = sw.Field(torch.from_numpy(f_ref.astype(np.float32)).cuda(), S.grid)
    for (mmin, eights, odd_first), xcd, pair in itertools.product(POLICIES, [True, False], [True, False]):
        S.set_tuning(mmin, eights, odd_first, xcd, pair)
        print(f"{name} {precision} policy {(mmin, eights, odd_first, xcd, pair)}", flush=True)
        f = sw.transform(coeffs, S).data.cpu().numpy()
        b = sw.transform(fin, S).data.cpu().numpy()
        assert rel(f, f_ref) < 1e-5
        assert rel(b, b_ref) < 1e-5
    S.close()


def test_autotune_choice_cached_and_consistent(sw, gpu):
    o, a, layers, fields = case(sw, "T127", seed=2)
    S = fresh_plan(sw, "T127")
    coeffs = to_dev(sw, a, S.spectrum)
    base = sw.transform(coeffs, S).data.cpu().numpy()
    t = S.autotune()
    assert t["tuned"] and t["mfma_min"] in (11, 17)
    tuned = sw.transform(coeffs, S).data.cpu().numpy()
    assert rel(tuned, base) < 2e-6
    # a second plan of the same size class takes the cached choice
    S2 = fresh_plan(sw, "T127")
    t2 = S2.autotune()
    assert t2 == t
    S.close()
    S2.close()


@pytest.mark.parametrize("precision", ["fp32", "fp16x3", "bf16x3"])
def test_legendre_launch_orders_bitwise_equal(sw, gpu, precision):
    """The four Legendre launch orders (XCD grouping, lightest tiles first) run the same
    tiles, so both transforms come out bitwise equal."""
    o, a, layers, fields = case(sw, "T255", seed=3)
    S = fresh_plan(sw, "T255", precision)
    coeffs = to_dev(sw, a, S.spectrum)
    outs = []
    for xcd in (False, True):
        for light in (False, True):
            S.set_tuning(17, False, True, xcd, True, legendre_light_first=light)
            t = S.tuning()
            assert t["legendre_xcd_groups"] == xcd and t["legendre_light_first"] == light
            f = sw.transform(coeffs, S)
            b = sw.transform(f, S)
            outs.append((f.data.cpu().numpy(), b.data.cpu().numpy()))
    for f, b in outs[1:]:
        assert np.array_equal(f, outs[0][0]) and np.array_equal(b, outs[0][1])
    S.close()
