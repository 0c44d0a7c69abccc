"""The allocation-free transform entry point (sht_transform_static, scratch in a static
device array) against sht_synthesis / sht_analysis through the C ABI: bitwise equal
results (the same kernels on the same inputs), on two calls in a row of which the second
is smaller, and the size limit rejected before any launch."""

import numpy as np
import pytest
import torch

from sht_cases import CONFIGS, oracle, rel

pytestmark = pytest.mark.gpu

SYN, ANA = 0, 1


def make_plan(sw, name, precision, nlayers):
    kind, n, lmax, mmax, _, _ = CONFIGS[name]
    grid = (sw.OctahedralGaussianGrid if kind == "octahedral" else sw.FullGaussianGrid)(n)
    return sw.SpectralTransform(sw.Spectrum(lmax, mmax), grid, nlayers=nlayers, precision=precision)


def static_call(sw, S, direction, inp, ld_in, out, ld_out, K, unscale=0):
    L = sw.lib()
    return L.sht_transform_static(S._plan, direction, inp.data_ptr(), ld_in, out.data_ptr(), ld_out, K, unscale,
                                  S._stream())


def regular_call(sw, S, direction, inp, ld_in, out, ld_out, K, unscale=0):
    L = sw.lib()
    if direction == SYN:
        return L.sht_synthesis(S._plan, inp.data_ptr(), ld_in, out.data_ptr(), ld_out, K, unscale, S._stream())
    return L.sht_analysis(S._plan, inp.data_ptr(), ld_in, out.data_ptr(), ld_out, K, S._stream())


@pytest.mark.parametrize("name, precision, K1, K2", [("T255", "fp16x3", 64, 5), ("T127", "fp32", 32, 3),
                                                     ("T31full", "bf16x3", 8, 1)])
def test_static_scratch_matches_regular_entry_points(sw, gpu, name, precision, K1, K2):
    print(f"static scratch case {name} {precision} K={K1} then K={K2}", flush=True)
    kind, n, lmax, mmax, _, _ = CONFIGS[name]
    S = make_plan(sw, name, precision, K1)
    nharm, npts = S.spectrum.nharm, S.grid.npoints
    o = oracle(kind, n, lmax, mmax)
    for K, seed, unscale in ((K1, 0, 0), (K2, 1, 1)):
        a = sw.seeded_spectrum(S.spectrum, K, seed=seed)
        coeffs = torch.from_numpy(a).to(torch.complex64).cuda()
        f_reg = torch.full((K, npts), float("nan"), device="cuda")
        f_sta = torch.full((K, npts), float("nan"), device="cuda")
        b_reg = torch.full((K, nharm), float("nan"), dtype=torch.complex64, device="cuda")
        b_sta = torch.full((K, nharm), float("nan"), dtype=torch.complex64, device="cuda")
        assert regular_call(sw, S, SYN, coeffs, nharm, f_reg, npts, K, unscale) == 0
        assert static_call(sw, S, SYN, coeffs, nharm, f_sta, npts, K, unscale) == 0, sw.lib().sht_last_error()
        assert regular_call(sw, S, ANA, f_reg, npts, b_reg, nharm, K) == 0
        assert static_call(sw, S, ANA, f_reg, npts, b_sta, nharm, K) == 0, sw.lib().sht_last_error()
        torch.cuda.synchronize()
        fr, fs = f_reg.cpu().numpy(), f_sta.cpu().numpy()
        br, bs = b_reg.cpu().numpy(), b_sta.cpu().numpy()
        assert np.isfinite(fs).all() and np.isfinite(bs).all()
        assert np.array_equal(fr, fs), ("synthesis", K)
        assert np.array_equal(br, bs), ("analysis", K)
        if not unscale:  # and both against the FP64 oracle, as the regular entry points are tested
            f_ref = o.synthesis(a)
            assert rel(fs.astype(np.float64), f_ref) < 1e-5
            assert rel(bs.astype(np.complex128), o.analysis(fs.astype(np.float64))) < 1e-5
    S.close()


def test_static_scratch_rejects_oversized_and_bad_calls(sw, gpu):
    import ctypes as C
    S = make_plan(sw, "T255", "fp32", 112)
    L = sw.lib()
    nharm, npts = S.spectrum.nharm, S.grid.npoints
    sizes = sw._lib.ShtSizes()
    assert L.sht_plan_sizes(C.byref(S._desc), C.byref(sizes)) == 0
    # T255: sum(nlon/2 + 1) = 38784 Fourier rows of 16 bytes per column -> K <= 108 in 64 MiB
    kmax = (64 << 20) // (sizes.nfreq_total * 16)
    assert kmax == 108
    coeffs = torch.zeros((112, nharm), dtype=torch.complex64, device="cuda")
    field = torch.zeros((112, npts), device="cuda")
    assert static_call(sw, S, SYN, coeffs, nharm, field, npts, kmax + 1) == sw._lib.SHT_ERR_DIM
    assert b"scratch" in L.sht_last_error()
    assert static_call(sw, S, SYN, coeffs, nharm, field, npts, 0) == sw._lib.SHT_ERR_DIM
    assert static_call(sw, S, ANA, field, npts, coeffs, nharm, 4, 1) == sw._lib.SHT_ERR_ARG  # unscale on analysis
    assert static_call(sw, S, 2, coeffs, nharm, field, npts, 4) == sw._lib.SHT_ERR_ARG      # bad direction
    # the largest K that fits still runs and matches the regular entry point
    a = sw.seeded_spectrum(S.spectrum, kmax, seed=7)
    c = torch.from_numpy(a).to(torch.complex64).cuda()
    f1 = torch.empty((kmax, npts), device="cuda")
    f2 = torch.empty((kmax, npts), device="cuda")
    assert regular_call(sw, S, SYN, c, nharm, f1, npts, kmax) == 0
    assert static_call(sw, S, SYN, c, nharm, f2, npts, kmax) == 0
    torch.cuda.synchronize()
    assert torch.equal(f1, f2)
    S.close()
