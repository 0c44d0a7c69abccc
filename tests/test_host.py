"""CPU tests of the product's host side: the C ABI library loads and exports every symbol
include/sht.h declares; host precompute (Gauss nodes, Λ, sizes) matches the oracle; the
Python mirror's sizes and synthetic inputs."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle.sht_oracle import Grid, OracleTransform, Spectrum as OSpectrum, gauss_legendre, legendre_lambda

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols(sw):
    header = open(os.path.join(REPO, "include", "sht.h")).read()
    declared = set(re.findall(r"\b(sht_[a-z_0-9]+)\s*\(", header))
    assert declared == set(sw._lib.EXPORTS)
    L = sw.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.sht_version().decode().startswith("libsht")


@pytest.mark.parametrize("N", [48, 192, 384])
def test_host_gauss_legendre(sw, N):
    x = np.zeros(N)
    w = np.zeros(N)
    assert sw.lib().sht_gauss_legendre(N, x.ctypes.data, w.ctypes.data) == 0
    xo, wo = gauss_legendre(N)
    assert np.abs(x - xo).max() < 2e-16
    assert np.max(np.abs(w - wo) / wo) < 1e-12


@pytest.mark.parametrize("lmax,mmax", [(33, 32), (257, 256)])
def test_host_legendre_column(sw, lmax, mmax):
    sp = OSpectrum(lmax, mmax)
    for x in [0.999, 0.7, 0.01, -0.4]:
        out = np.zeros(sp.nharm)
        assert sw.lib().sht_legendre_column(lmax, mmax, x, out.ctypes.data) == 0
        ref = legendre_lambda(lmax, mmax, x)
        scale = np.abs(ref).max()
        assert np.abs(out - ref).max() < 1e-13 * scale


@pytest.mark.parametrize("grid,n,lmax,mmax", [(0, 24, 33, 32), (0, 96, 129, 128), (0, 192, 257, 256),
                                              (1, 24, 33, 32)])
def test_plan_sizes(sw, grid, n, lmax, mmax):
    d = sw._lib.ShtDesc(grid, n, lmax, mmax, 8, 0, 0, 0)
    s = sw._lib.ShtSizes()
    assert sw.lib().sht_plan_sizes(C.byref(d), C.byref(s)) == 0
    og = Grid("octahedral" if grid == 0 else "full", n)
    assert s.npoints == og.npoints
    assert s.nharm == OSpectrum(lmax, mmax).nharm
    assert s.nlat == 2 * n
    assert s.nlon_max == og.nlons.max()
    assert s.nfreq_total == int((og.nlons_half // 2 + 1).sum())
    pg = (sw.OctahedralGaussianGrid if grid == 0 else sw.FullGaussianGrid)(n)
    assert pg.npoints == og.npoints and list(pg.nlons) == list(og.nlons)


def test_plan_rejects_bad_descriptors(sw):
    s = sw._lib.ShtSizes()
    for d in [sw._lib.ShtDesc(7, 24, 33, 32, 8, 0, 0, 0),   # unknown grid
              sw._lib.ShtDesc(0, 24, 31, 32, 8, 0, 0, 0),   # lmax < mmax
              sw._lib.ShtDesc(0, 24, 33, 32, 0, 0, 0, 0)]:  # no layers
        assert sw.lib().sht_plan_sizes(C.byref(d), C.byref(s)) != 0
        assert len(sw.lib().sht_last_error()) > 0


def test_resolution_helpers(sw):
    assert [sw.get_nlat_half(t) for t in (32, 128, 256)] == [24, 96, 192]
    assert sw.get_truncation(24) == 32
    sp = sw.Spectrum.from_truncation(255 + 1, one_more_degree=True)
    assert (sp.lmax, sp.mmax, sp.nharm) == (257, 256, 33152)
    assert sp.lm2i(0, 0) == 0 and sp.lm2i(256, 0) == 256 and sp.lm2i(1, 1) == 257
    rings = sw.OctahedralGaussianGrid(24).eachring()
    assert rings[0] == range(0, 20) and rings[-1].stop == sw.OctahedralGaussianGrid(24).npoints


def test_seeded_spectrum(sw):
    sp = sw.Spectrum(33, 32)
    a = sw.seeded_spectrum(sp, 3, seed=4)
    b = sw.seeded_spectrum(sp, 3, seed=4)
    assert np.array_equal(a, b)
    ls, ms = sp.degrees_orders()
    assert np.all(a[:, ms == 0].imag == 0)
    # amplitude envelope (l+1)^-1.5 with unit-variance normal draws
    z = a * (ls + 1.0) ** 1.5
    assert abs(np.std(z[:, ms > 0].real) - 1) < 0.1
    assert not np.array_equal(a[0], a[1])
    c = sw.seeded_spectrum(sp, 3, seed=5)
    assert not np.array_equal(a, c)


def test_no_cpu_fallback_without_gpu(sw):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(sw.SHTError):
        sw.SpectralTransform(sw.Spectrum(33, 32), sw.OctahedralGaussianGrid(24), nlayers=2)


def test_tune_cache_file_round_trip(sw, tmp_path):
    """Autotuner choices persist as one JSON object per line: a loaded file saves back with
    the same choices; a malformed line is reported while the valid ones still load."""
    import ctypes as C
    import json
    L = sw.lib()
    keys = ["grid", "nlat_half", "lmax", "mmax", "nlayers", "precision", "K", "mfma_min", "eights", "odd_first",
            "leg_xcd", "fft_pair", "fourier", "dft_cg"]
    rows = [dict(zip(keys, [0, 192, 257, 256, 64, 1, 64, 17, 0, 1, 1, 1, 2, 8])),
            dict(zip(keys, [1, 24, 33, 32, 8, 0, 8, 11, 1, 0, 0, 1, 0, 4]))]
    src = tmp_path / "tune.jsonl"
    src.write_text("".join(json.dumps(r) + "\n" for r in rows))
    n = C.c_int32()
    assert L.sht_tune_cache_load(str(src).encode(), C.byref(n)) == 0 and n.value == 2
    out = tmp_path / "saved.jsonl"
    assert L.sht_tune_cache_save(str(out).encode()) == 0
    saved = [json.loads(x) for x in out.read_text().splitlines() if x.strip()]
    for r in rows:
        assert r in saved
    bad = tmp_path / "bad.jsonl"
    r3 = dict(rows[0], K=32)
    bad.write_text(json.dumps(r3) + "\n" + '{"grid": 0, "K": 5}\n' + json.dumps(dict(rows[0], dft_cg=3)) + "\n")
    assert L.sht_tune_cache_load(str(bad).encode(), C.byref(n)) == 1  # SHT_ERR_ARG
    assert n.value == 1
    assert b"malformed" in L.sht_last_error()
    assert L.sht_tune_cache_load(b"/nonexistent/dir/x.jsonl", None) == 1
