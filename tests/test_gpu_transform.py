"""GPU parity of the HIP engine against the FP64 oracle and the FP32 C restatement of the
reference CPU path, through the C ABI (libsht.so) via the Python mirror of transform!."""
import math

import numpy as np
import pytest
import torch

from oracle.cref import CRefTransform
from oracle.sht_oracle import uv_from_vordiv
from sht_cases import CONFIGS, oracle, per_field_band_errors, rel

pytestmark = pytest.mark.gpu

_PLANS = {}


def plan(sw, name, precision, nlayers=None):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    K = nlayers or layers * fields
    key = (name, precision, K)
    if key not in _PLANS:
        grid = (sw.OctahedralGaussianGrid if kind == "octahedral" else sw.FullGaussianGrid)(n)
        _PLANS[key] = sw.SpectralTransform(sw.Spectrum(lmax, mmax), grid, nlayers=K, precision=precision)
    return _PLANS[key]


def to_dev(sw, a, spectrum):
    return sw.LowerTriangularArray(torch.from_numpy(np.ascontiguousarray(a)).to(torch.complex64).cuda(), spectrum)


def case(sw, name, seed=0):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    o = oracle(kind, n, lmax, mmax)
    a = sw.seeded_spectrum(sw.Spectrum(lmax, mmax), layers * fields, seed=seed)
    return o, a, layers, fields


# tolerance of each precision mode, relative L2 per field (and per degree band for round trips)
TOL = {"fp32": 1e-5, "fp16x3": 1e-5, "bf16x3": 1e-5}


@pytest.mark.parametrize("name", ["T31", "T127", "T255", "T31full"])
@pytest.mark.parametrize("precision", ["fp32", "fp16x3", "bf16x3"])
def test_synthesis_analysis_roundtrip_vs_fp64(sw, gpu, name, precision):
    o, a, layers, fields = case(sw, name)
    S = plan(sw, name, precision)
    print(f"case {name} {precision} K={layers * fields}", flush=True)
    coeffs = to_dev(sw, a, S.spectrum)
    field = sw.transform(coeffs, S)
    back = sw.transform(field, S)
    # analysis of the exact (FP64 oracle) grid
    f_ref = o.synthesis(a)
    fin = sw.Field(torch.from_numpy(f_ref.astype(np.float32)).cuda(), S.grid)
    ana = sw.transform(fin, S)
    torch.cuda.synchronize()
    f_gpu = field.data.cpu().numpy().astype(np.float64)
    b_gpu = back.data.cpu().numpy().astype(np.complex128)
    ana_gpu = ana.data.cpu().numpy().astype(np.complex128)
    b_ref = o.analysis(f_ref)
    tol = TOL[precision]
    for f in range(fields):
        cols = slice(f * layers, (f + 1) * layers)
        assert rel(f_gpu[cols], f_ref[cols]) < tol, ("synthesis", f)
        assert rel(ana_gpu[cols], b_ref[cols]) < tol, ("analysis", f)
    errs = per_field_band_errors(b_gpu, b_ref, o.spectrum.l_of, layers, fields)
    worst = max(errs.values())
    print(f"{name} {precision}: round-trip worst field/band rel err {worst:.2e}", flush=True)
    assert worst < tol, errs
    assert np.isfinite(f_gpu).all() and np.isfinite(b_gpu).all()


@pytest.mark.parametrize("name", ["T31", "T127", "T255"])
def test_fp32_vs_cpu_fp32_restatement(sw, gpu, name):
    o, a, layers, fields = case(sw, name, seed=3)
    S = plan(sw, name, "fp32")
    c = CRefTransform(o)
    f_cpu = c.synthesis(a, np.float32)
    b_cpu = c.analysis(f_cpu, np.float32)
    coeffs = to_dev(sw, a, S.spectrum)
    field = sw.transform(coeffs, S)
    fin = sw.Field(torch.from_numpy(f_cpu).cuda(), S.grid)
    back = sw.transform(fin, S)
    torch.cuda.synchronize()
    print(f"{name}: FP32 GPU vs FP32 C restatement", flush=True)
    assert rel(field.data.cpu().numpy(), f_cpu) < 1e-5
    assert rel(back.data.cpu().numpy(), b_cpu) < 1e-5
    # per field and degree band of the analysis, as the round-trip criterion is stated
    errs = per_field_band_errors(back.data.cpu().numpy().astype(np.complex128), b_cpu.astype(np.complex128),
                                 o.spectrum.l_of, layers, fields)
    assert max(errs.values()) < 1e-5, errs


def test_two_streams_share_one_plan(sw, gpu):
    """Round trips on two streams at once on one plan (each stream has its own Fourier
    scratch) equal the same round trips run one after the other, bit for bit."""
    S = plan(sw, "T127", "fp16x3")
    K = 32
    ins = [to_dev(sw, sw.seeded_spectrum(S.spectrum, K, seed=40 + i), S.spectrum) for i in range(2)]
    refs = [sw.transform(sw.transform(c, S), S).data.clone() for c in ins]
    fields = [sw.Field(torch.empty((K, S.grid.npoints), device="cuda"), S.grid) for _ in range(2)]
    outs = [sw.LowerTriangularArray.zeros(S.spectrum, K, device="cuda") for _ in range(2)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    for rep in range(4):
        for i in (0, 1):
            with torch.cuda.stream(streams[i]):
                sw.transform_(fields[i], ins[i], S)
                sw.transform_(outs[i], fields[i], S)
    torch.cuda.synchronize()
    for i in (0, 1):
        assert torch.equal(outs[i].data, refs[i]), i


def test_y00_constant_grid(sw, gpu):
    S = plan(sw, "T127", "fp32")
    a = np.zeros((3, S.spectrum.nharm), complex)
    a[:, 0] = 2 * math.sqrt(math.pi) * np.array([1.0, 2.0, -0.5])
    f = sw.transform(to_dev(sw, a, S.spectrum), S)
    torch.cuda.synchronize()
    g = f.data.cpu().numpy()
    for k, c in enumerate([1.0, 2.0, -0.5]):
        assert np.abs(g[k] - c).max() < 2e-6 * abs(c)


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_nan_poisoned_outputs_fully_written(sw, gpu, precision):
    # any output element the kernels miss stays NaN (SpeedyWeather/test/GPU/spectral_transform.jl:235-295)
    o, a, layers, fields = case(sw, "T127", seed=1)
    K = 5  # odd: partial column groups in every kernel
    S = plan(sw, "T127", precision)
    coeffs = to_dev(sw, a[:K], S.spectrum)
    field = sw.Field(torch.full((K, S.grid.npoints), float("nan"), device="cuda"), S.grid)
    sw.transform_(field, coeffs, S)
    back = sw.LowerTriangularArray(torch.full((K, S.spectrum.nharm), complex("nan+nanj"), dtype=torch.complex64,
                                              device="cuda"), S.spectrum)
    sw.transform_(back, field, S)
    torch.cuda.synchronize()
    assert torch.isfinite(field.data).all()
    assert torch.isfinite(torch.view_as_real(back.data)).all()
    f_ref = o.synthesis(a[:K])
    assert rel(field.data.cpu().numpy(), f_ref) < TOL[precision]


def test_k_subsets_and_strided_columns(sw, gpu):
    o, a, layers, fields = case(sw, "T31", seed=2)
    S = plan(sw, "T31", "fp32")  # nlayers 8
    f_ref = o.synthesis(a)
    b_ref = o.analysis(f_ref)
    for K in (1, 3, 8):
        c = to_dev(sw, a[:K], S.spectrum)
        f = sw.transform(c, S)
        torch.cuda.synchronize()
        assert rel(f.data.cpu().numpy(), f_ref[:K]) < 1e-5
    # 1-D (single layer) arrays
    f1 = sw.transform(sw.LowerTriangularArray(c.data[0].clone(), S.spectrum), S)
    torch.cuda.synchronize()
    assert f1.data.dim() == 1 and rel(f1.data.cpu().numpy(), f_ref[0]) < 1e-5
    # columns with a stride larger than the column length (views into a wider buffer)
    wide_c = torch.zeros((8, S.spectrum.nharm + 37), dtype=torch.complex64, device="cuda")
    wide_c[:, :S.spectrum.nharm] = torch.from_numpy(a).to(torch.complex64)
    wide_f = torch.full((8, S.grid.npoints + 101), 7.0, device="cuda")
    cv = sw.LowerTriangularArray(wide_c[:, :S.spectrum.nharm], S.spectrum)
    fv = sw.Field(wide_f[:, :S.grid.npoints], S.grid)
    sw.transform_(fv, cv, S)
    back = sw.LowerTriangularArray(torch.zeros((8, S.spectrum.nharm + 5), dtype=torch.complex64,
                                               device="cuda")[:, :S.spectrum.nharm], S.spectrum)
    sw.transform_(back, fv, S)
    torch.cuda.synchronize()
    assert rel(wide_f[:, :S.grid.npoints].cpu().numpy(), f_ref) < 1e-5
    assert torch.all(wide_f[:, S.grid.npoints:] == 7.0)  # padding untouched
    assert rel(back.data.cpu().numpy(), b_ref) < 1e-5


@pytest.mark.parametrize("kind, n, lmax, mmax", [("octahedral", 20, 29, 28), ("full", 12, 17, 16)])
@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_ring_counts_not_a_multiple_of_the_xcds(sw, gpu, kind, n, lmax, mmax, precision):
    """Ring counts that are not multiples of 8: the FFT synthesis puts the column groups of ring
    rank r on XCD r mod 8 and pads the ranks to a multiple of 8 (the padding blocks exit), with
    1, 3 and 5 columns (partial column groups); both directions and the round trip against the
    FP64 oracle, and the same fields with the column groups in rank order (fft_pairs off)."""
    print(f"ring count case {kind} n={n} {precision}", flush=True)
    o = oracle(kind, n, lmax, mmax)
    grid = (sw.OctahedralGaussianGrid if kind == "octahedral" else sw.FullGaussianGrid)(n)
    sp = sw.Spectrum(lmax, mmax)
    S = sw.SpectralTransform(sp, grid, nlayers=5, precision=precision)
    a = sw.seeded_spectrum(sp, 5, seed=31)
    f_ref = o.synthesis(a)
    b_ref = o.analysis(f_ref)
    for K in (1, 3, 5):
        f = sw.transform(to_dev(sw, a[:K], sp), S)
        b = sw.transform(f, S)
        torch.cuda.synchronize()
        assert rel(f.data.cpu().numpy().reshape(K, -1), f_ref[:K]) < TOL[precision], K
        assert rel(b.data.cpu().numpy().reshape(K, -1), b_ref[:K]) < TOL[precision], K
    f_grouped = sw.transform(to_dev(sw, a, sp), S).data.clone()
    tuning = S.tuning()
    S.set_tuning(tuning["mfma_min"], tuning["eights"], tuning["odd_first"], tuning["legendre_xcd_groups"],
                 False, tuning["fourier"], tuning["dft_cols"], tuning["legendre_light_first"])
    f_plain = sw.transform(to_dev(sw, a, sp), S).data
    torch.cuda.synchronize()
    assert torch.equal(f_grouped, f_plain)  # the launch order changes no arithmetic
    S.close()


@pytest.mark.parametrize("offset, extra", [(0, 4), (4, 4), (1, 4), (2, 6)])
def test_grid_views_aligned_and_unaligned(sw, gpu, offset, extra):
    """Grid views at any float offset and column stride (aligned or not to 8 or 16 bytes)
    reproduce the oracle and leave the rest of the buffer untouched."""
    o, a, layers, fields = case(sw, "T31", seed=5)
    S = plan(sw, "T31", "fp32")
    f_ref = o.synthesis(a)
    b_ref = o.analysis(f_ref)
    npts = S.grid.npoints
    wide = torch.full((8, npts + extra), 7.0, device="cuda")
    fv = sw.Field(wide[:, offset:offset + npts], S.grid)
    sw.transform_(fv, to_dev(sw, a, S.spectrum), S)
    back = sw.transform(fv, S)
    torch.cuda.synchronize()
    assert rel(fv.data.cpu().numpy(), f_ref) < 1e-5
    assert rel(back.data.cpu().numpy(), b_ref) < 1e-5
    outside = torch.cat([wide[:, :offset], wide[:, offset + npts:]], dim=1)
    assert torch.all(outside == 7.0)


def test_unscale_coslat(sw, gpu):
    o, a, layers, fields = case(sw, "T127", seed=4)
    S = plan(sw, "T127", "fp32")
    f = sw.transform(to_dev(sw, a, S.spectrum), S, unscale_coslat=True)
    torch.cuda.synchronize()
    f_ref = o.synthesis(a, unscale_coslat=True)
    assert rel(f.data.cpu().numpy(), f_ref) < 1e-5


def test_dimension_mismatch(sw, gpu):
    S = plan(sw, "T31", "fp32")  # nlayers 8
    sp = S.spectrum
    with pytest.raises(sw.DimensionMismatch):  # more columns than the plan holds
        sw.transform(sw.LowerTriangularArray(torch.zeros((9, sp.nharm), dtype=torch.complex64,
                                                         device="cuda"), sp), S)
    with pytest.raises(sw.DimensionMismatch):  # other spectrum
        sp2 = sw.Spectrum(32, 32)
        sw.transform(sw.LowerTriangularArray(torch.zeros((2, sp2.nharm), dtype=torch.complex64,
                                                         device="cuda"), sp2), S)
    with pytest.raises(sw.DimensionMismatch):  # other grid
        g2 = sw.FullGaussianGrid(24)
        sw.transform(sw.Field(torch.zeros((2, g2.npoints), device="cuda"), g2), S)
    with pytest.raises(TypeError):  # grid data must be Float32 like the transform
        sw.transform(sw.Field(torch.zeros((2, S.grid.npoints), dtype=torch.float64, device="cuda"), S.grid), S)
    with pytest.raises(ValueError):  # no CPU path
        sw.transform(sw.LowerTriangularArray(torch.zeros((2, sp.nharm), dtype=torch.complex64), sp), S)
    # the C ABI itself rejects K above capacity
    import ctypes as C
    rc = sw.lib().sht_synthesis(S._plan, C.c_void_p(1), sp.nharm, C.c_void_p(1), S.grid.npoints, 9, 0, None)
    assert rc == 2


def test_complex128_coefficients_accepted(sw, gpu):
    # spec->grid accepts Complex{Float64} coefficients into a Float32 transform
    # (SpeedyTransforms/test/spectral_transform.jl:183-219)
    o, a, layers, fields = case(sw, "T31", seed=6)
    S = plan(sw, "T31", "fp32")
    c = sw.LowerTriangularArray(torch.from_numpy(a).cuda(), S.spectrum)
    f = sw.transform(c, S)
    torch.cuda.synchronize()
    assert rel(f.data.cpu().numpy(), o.synthesis(a)) < 1e-5


def test_deterministic(sw, gpu):
    o, a, layers, fields = case(sw, "T255", seed=0)
    S = plan(sw, "T255", "fp32")
    c = to_dev(sw, a, S.spectrum)
    f1 = sw.transform(c, S)
    b1 = sw.transform(f1, S)
    f2 = sw.transform(c, S)
    b2 = sw.transform(f2, S)
    torch.cuda.synchronize()
    assert torch.equal(f1.data, f2.data) and torch.equal(b1.data, b2.data)


@pytest.mark.parametrize("name", ["T31", "T127"])
def test_uv_from_vordiv(sw, gpu, name):
    kind, n, lmax, mmax, layers, fields = CONFIGS[name]
    sp = sw.Spectrum(lmax, mmax)
    S = plan(sw, name, "fp32")
    K = 6
    vor = sw.seeded_spectrum(sp, K, seed=11)
    div = sw.seeded_spectrum(sp, K, seed=12)
    U_ref, V_ref = uv_from_vordiv(oracle(kind, n, lmax, mmax).spectrum, vor, div, radius=2.5)
    U = sw.LowerTriangularArray.zeros(sp, K, device="cuda")
    V = sw.LowerTriangularArray.zeros(sp, K, device="cuda")
    sw.UV_from_vordiv_(U, V, to_dev(sw, vor, sp), to_dev(sw, div, sp), S, radius=2.5)
    torch.cuda.synchronize()
    assert rel(U.data.cpu().numpy(), U_ref) < 1e-6
    assert rel(V.data.cpu().numpy(), V_ref) < 1e-6
    # vector path into grid space: u cos(lat) then unscale -> u (spectral_grid transform.jl:145,162-166)
    u = sw.transform(U, S, unscale_coslat=True)
    torch.cuda.synchronize()
    u_ref = oracle(kind, n, lmax, mmax).synthesis(U_ref, unscale_coslat=True)
    assert rel(u.data.cpu().numpy(), u_ref) < 1e-5
