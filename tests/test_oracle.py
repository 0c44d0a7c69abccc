"""Pins of the FP64 oracle (oracle/sht_oracle.py) and of its C restatement against
closed forms, library routines, invariants and the reference's own test properties."""
import math

import numpy as np
import pytest
import scipy.special

from oracle.cref import CRefTransform, fft64
from oracle.sht_oracle import divergence as o_divergence, gradient as o_gradient, gradient_factors
from oracle.sht_oracle import laplace as o_laplace
from oracle.sht_oracle import (Grid, OracleTransform, Spectrum, gauss_legendre, get_nlat_half,
                               get_truncation, legendre_lambda, roundup_fft, uv_from_vordiv)


# ------------------------------------------------------------- sizes (aliasing.jl)
@pytest.mark.parametrize("dealiasing", [2, 3])
def test_resolution_kat(dealiasing):
    # SpeedyTransforms/test/spectral_transform.jl:1-17
    for i in range(5, 11):
        trunc = 2 ** i
        nlon = (dealiasing + 1) * 2 ** i
        n = get_nlat_half(trunc, dealiasing)
        assert (nlon, nlon // 2) == (4 * n, 2 * n)
        assert get_truncation(n, dealiasing) == trunc


def test_roundup_fft_kat():
    # SpeedyTransforms/test/resolutions.jl:9-19
    for i in range(1, 11):
        assert roundup_fft(2 ** i) == 2 ** i
        assert roundup_fft(3 * 2 ** i) == 3 * 2 ** i
        assert roundup_fft(5 * 2 ** i) == 5 * 2 ** i
    for i in [7, 97, 331, 997]:
        assert roundup_fft(i) >= i


def test_config_geometry():
    # T31/T127/T255 (1-based mmax 32/128/256) -> nlat_half 24/96/192; npoints 4n^2+36n
    for mmax, n in [(32, 24), (128, 96), (256, 192)]:
        assert get_nlat_half(mmax) == n
        g = Grid("octahedral", n)
        assert g.npoints == 4 * n * n + 36 * n
    assert Spectrum(257, 256).nharm == 33152 and Spectrum(33, 32).nharm == 560


# ------------------------------------------------------------- Gauss-Legendre
@pytest.mark.parametrize("N", [2, 7, 48, 192, 384])
def test_gauss_legendre(N):
    x, w = gauss_legendre(N)
    xr, wr = np.polynomial.legendre.leggauss(N)
    assert np.allclose(x, xr, atol=1e-14, rtol=0)
    assert np.allclose(w, wr, atol=0, rtol=1e-9)  # leggauss loses ~5e-11 relative at the poles
    # extended-precision Newton refinement of the same nodes pins the weights to 1e-11
    xl = x.astype(np.longdouble)
    for _ in range(3):
        p0, p1 = np.ones_like(xl), xl.copy()
        for k in range(2, N + 1):
            p0, p1 = p1, ((2 * k - 1) * xl * p1 - (k - 1) * p0) / k
        dp = N * (xl * p1 - p0) / (xl * xl - 1)
        xl = xl - p1 / dp
    wl = 2 / ((1 - xl * xl) * dp * dp)
    assert float(np.max(np.abs(w - wl) / wl)) < 1e-11
    assert abs(w.sum() - 2) < 1e-13
    assert np.allclose(x, -x[::-1], atol=1e-15)
    for p in range(0, 2 * N):  # exact for degree <= 2N-1
        exact = 0.0 if p % 2 else 2.0 / (p + 1)
        assert abs(np.dot(w, x ** p) - exact) < 1e-13


# ------------------------------------------------------------- Legendre Λ
def test_legendre_closed_forms():
    sp = Spectrum(4, 3)
    for x in [-0.9, -0.3, 0.0, 0.2, 0.77, 0.999]:
        s = math.sqrt(1 - x * x)
        lam = legendre_lambda(4, 3, x)
        v = lambda l, m: lam[sp.index(l, m)]
        pi = math.pi
        assert v(0, 0) == pytest.approx(1 / (2 * math.sqrt(pi)), rel=1e-15)
        assert v(1, 0) == pytest.approx(math.sqrt(3 / (4 * pi)) * x, rel=1e-14, abs=1e-16)
        assert v(1, 1) == pytest.approx(-math.sqrt(3 / (8 * pi)) * s, rel=1e-14)
        assert v(2, 0) == pytest.approx(math.sqrt(5 / (16 * pi)) * (3 * x * x - 1), rel=1e-13, abs=1e-16)
        assert v(2, 1) == pytest.approx(-math.sqrt(15 / (8 * pi)) * x * s, rel=1e-13, abs=1e-16)
        assert v(2, 2) == pytest.approx(math.sqrt(15 / (32 * pi)) * s * s, rel=1e-13)


def test_legendre_vs_scipy():
    lmax = mmax = 120
    sp = Spectrum(lmax, mmax)
    xs = np.array([-0.95, -0.4, 0.05, 0.6, 0.98])
    lam = legendre_lambda(lmax, mmax, xs)
    theta = np.arccos(xs)
    for m in [0, 1, 2, 5, 17, 60, 119]:
        for l in range(m, lmax, 7):
            y = scipy.special.sph_harm_y(l, m, theta, 0.0).real
            np.testing.assert_allclose(lam[sp.index(l, m)], y, rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("n", [24, 96])
def test_legendre_discrete_orthonormality(n):
    # 2 pi sum_j w_j λ_l^m λ_l'^m = δ_ll' over both hemispheres (exact for l + l' <= 4n-1)
    lmax, mmax = 2 * n - 1, 2 * n - 1
    x, w = gauss_legendre(2 * n)
    lam = legendre_lambda(lmax, mmax, -x)
    sp = Spectrum(lmax, mmax)
    for m in [0, 1, 7, mmax - 1]:
        o = sp.offsets[m]
        L = lam[o:o + lmax - m]
        G = 2 * np.pi * (L * w) @ L.T
        assert np.abs(G - np.eye(lmax - m)).max() < 1e-12


def test_legendre_recursion_residual():
    # x λ_l^m = ε_{l+1}^m λ_{l+1}^m + ε_l^m λ_{l-1}^m (orthonormal three-term identity)
    lmax = mmax = 257
    sp = Spectrum(lmax, mmax)
    xs = np.array([0.01, 0.3, 0.9])
    lam = legendre_lambda(lmax, mmax, xs)
    eps = lambda l, m: math.sqrt((l * l - m * m) / (4.0 * l * l - 1))
    worst = 0.0
    for m in [0, 3, 100, 200]:
        for l in range(m + 1, lmax - 1):
            lhs = xs * lam[sp.index(l, m)]
            rhs = eps(l + 1, m) * lam[sp.index(l + 1, m)] + eps(l, m) * lam[sp.index(l - 1, m)]
            scale = np.abs(lam[sp.offsets[m]:sp.offsets[m] + lmax - m]).max(axis=0) + 1e-300
            worst = max(worst, float(np.max(np.abs(lhs - rhs) / scale)))
    assert worst < 1e-12


# ------------------------------------------------------------- FFT semantics
@pytest.mark.parametrize("N", [20, 24, 96, 772])
def test_rfft_brfft_semantics_bruteforce(N):
    rng = np.random.default_rng(N)
    f = rng.standard_normal(N)
    n = np.arange(N)
    nf = N // 2 + 1
    F = np.array([np.sum(f * np.exp(-2j * np.pi * m * n / N)) for m in range(nf)])
    np.testing.assert_allclose(np.fft.rfft(f), F, atol=1e-10)
    # brfft: Re X0 + 2 sum Re(X_m e^{+}) + Re X_{N/2} (-1)^n, imaginary DC/Nyquist ignored
    X = rng.standard_normal(nf) + 1j * rng.standard_normal(nf)
    ref = X[0].real + (X[nf - 1].real * (-1.0) ** n)
    for m in range(1, nf - 1):
        ref = ref + 2 * (X[m] * np.exp(2j * np.pi * m * n / N)).real
    np.testing.assert_allclose(np.fft.irfft(X, n=N) * N, ref, atol=1e-9)


@pytest.mark.parametrize("N", [20, 28, 388, 772, 776])
def test_cref_fft_vs_numpy(N):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    np.testing.assert_allclose(fft64(x, -1), np.fft.fft(x), atol=1e-11 * N)
    np.testing.assert_allclose(fft64(x, +1), np.fft.ifft(x) * N, atol=1e-11 * N)


# ------------------------------------------------------------- whole transform
@pytest.mark.parametrize("kind,n,trunc", [("octahedral", 24, 32), ("full", 24, 32), ("octahedral", 48, 64)])
def test_y00_constant(kind, n, trunc):
    # SpeedyTransforms/test/spectral_transform.jl:38-69
    sp = Spectrum(trunc + 1, trunc)
    o = OracleTransform(sp, Grid(kind, n))
    a = np.zeros((1, sp.nharm), complex)
    a[0, 0] = o.norm_sphere
    f = o.synthesis(a)
    assert np.abs(f - 1).max() < 1e-13


def test_single_harmonics_roundtrip():
    # SpeedyTransforms/test/spectral_transform.jl:71-102 (l <= 3, atol 100 eps)
    sp = Spectrum(33, 32)
    o = OracleTransform(sp, Grid("octahedral", 24))
    for l in range(4):
        for m in range(l + 1):
            for val in ([1.0, 1j] if m > 0 else [1.0]):
                a = np.zeros((1, sp.nharm), complex)
                a[0, sp.index(l, m)] = val
                b = o.analysis(o.synthesis(a))
                assert np.abs(b - a).max() < 100 * np.finfo(float).eps


@pytest.mark.parametrize("kind", ["octahedral", "full"])
@pytest.mark.parametrize("trunc", [43, 64])
def test_roundtrip_cubic_dealiasing(kind, trunc):
    # SpeedyTransforms/test/spectral_transform.jl:326-373: FP64, dealiasing 3, 8 layers,
    # grid -> spec -> grid rtol 1e-13
    sp = Spectrum(trunc, trunc)
    n = get_nlat_half(trunc, 3)
    o = OracleTransform(sp, Grid(kind, n))
    rng = np.random.default_rng(trunc)
    spec = rng.standard_normal((8, sp.nharm)) + 1j * rng.standard_normal((8, sp.nharm))
    grid = o.synthesis(spec)
    grid2 = o.synthesis(o.analysis(grid))
    assert np.linalg.norm(grid2 - grid) / np.linalg.norm(grid) < 1e-13


def test_quadratic_dealiasing_error_floor():
    # at the configs' quadratic dealiasing the polar rings drop orders: the FP64 spec->grid->spec
    # round trip is NOT the identity; it stays at ~1e-10 at T31 (input spectrum (l+1)^-1.5)
    sp = Spectrum(33, 32)
    o = OracleTransform(sp, Grid("octahedral", 24))
    rng = np.random.default_rng(0)
    ls = sp.l_of
    a = (rng.standard_normal((4, sp.nharm)) + 1j * rng.standard_normal((4, sp.nharm))) * (ls + 1.0) ** -1.5
    a[:, sp.m_of == 0] = a[:, sp.m_of == 0].real
    b = o.analysis(o.synthesis(a))
    err = np.linalg.norm(b - a) / np.linalg.norm(a)
    assert 1e-13 < err < 1e-8


def test_linearity_and_parseval():
    sp = Spectrum(33, 32)
    o = OracleTransform(sp, Grid("octahedral", 24))
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, sp.nharm)) + 1j * rng.standard_normal((2, sp.nharm))
    a[:, sp.m_of == 0] = a[:, sp.m_of == 0].real
    f0, f1 = o.synthesis(a[:1]), o.synthesis(a[1:])
    np.testing.assert_allclose(o.synthesis(2 * a[:1] - 3 * a[1:]), 2 * f0 - 3 * f1, atol=1e-12)
    # Parseval: sum dOmega f^2 = sum (2 - delta_m0)|a|^2 for a band-limited field
    lo = a.copy()
    lo[:, sp.l_of > 20] = 0
    f = o.synthesis(lo[:1])[0]
    dO = np.repeat(o.solid_angles, o.grid.nlons)
    lhs = np.sum(dO * f * f)
    rhs = np.sum((2 - (sp.m_of == 0)) * np.abs(lo[0]) ** 2)
    assert abs(lhs - rhs) / rhs < 1e-12


# ------------------------------------------------------------- C restatement
@pytest.mark.parametrize("kind,n,lmax,mmax", [("octahedral", 24, 33, 32), ("full", 24, 33, 32),
                                               ("octahedral", 96, 129, 128)])
def test_cref_matches_numpy(kind, n, lmax, mmax):
    sp = Spectrum(lmax, mmax)
    o = OracleTransform(sp, Grid(kind, n))
    c = CRefTransform(o)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((3, sp.nharm)) + 1j * rng.standard_normal((3, sp.nharm))
    a[:, sp.m_of == 0] = a[:, sp.m_of == 0].real
    f = o.synthesis(a)
    b = o.analysis(f)
    for unscale in (False, True):
        fu = o.synthesis(a, unscale_coslat=unscale)
        assert np.abs(c.synthesis(a, np.float64, unscale_coslat=unscale) - fu).max() < 1e-12 * np.abs(fu).max()
    assert np.abs(c.analysis(f, np.float64) - b).max() < 1e-12 * np.abs(b).max()
    f32 = c.synthesis(a, np.float32)
    b32 = c.analysis(f.astype(np.float32), np.float32)
    assert np.linalg.norm(f32 - f) / np.linalg.norm(f) < 2e-6
    assert np.linalg.norm(b32 - b) / np.linalg.norm(b) < 2e-6


# ------------------------------------------------------------- vector path
def _eval_field(sp, a, lat, lon):
    """f(lat, lon) = sum_l a_l0 λ + 2 Re sum_{m>0} a_lm λ e^{i m lon} at points."""
    lam = legendre_lambda(sp.lmax, sp.mmax, np.sin(lat))  # cos(colat) = sin(lat)
    out = np.zeros(lat.shape)
    for i in range(sp.nharm):
        m = sp.m_of[i]
        term = a[i] * lam[i] * np.exp(1j * m * lon)
        out += term.real if m == 0 else 2 * term.real
    return out


def test_uv_from_vordiv_vs_streamfunction():
    # U = -cos(lat) dψ/dlat + dχ/dlon,  V = dψ/dlon + cos(lat) dχ/dlat with ∇²ψ = ζ, ∇²χ = D
    sp = Spectrum(12, 11)
    rng = np.random.default_rng(7)
    vor = rng.standard_normal(sp.nharm) + 1j * rng.standard_normal(sp.nharm)
    div = rng.standard_normal(sp.nharm) + 1j * rng.standard_normal(sp.nharm)
    for x in (vor, div):
        x[sp.m_of == 0] = x[sp.m_of == 0].real
        x[sp.l_of == sp.lmax - 1] = 0  # the extra degree carries no vorticity/divergence
        x[0] = 0
    U, V = uv_from_vordiv(sp, vor, div)
    inv = np.zeros(sp.nharm)
    inv[sp.l_of > 0] = -1.0 / (sp.l_of[sp.l_of > 0] * (sp.l_of[sp.l_of > 0] + 1.0))
    psi, chi = vor * inv, div * inv
    lat = np.array([-1.1, -0.3, 0.2, 0.9, 1.3])
    lon = np.array([0.3, 2.0, 4.1, 5.5, 1.0])
    h = 1e-5
    dpsi_dlat = (_eval_field(sp, psi, lat + h, lon) - _eval_field(sp, psi, lat - h, lon)) / (2 * h)
    dchi_dlat = (_eval_field(sp, chi, lat + h, lon) - _eval_field(sp, chi, lat - h, lon)) / (2 * h)
    dpsi_dlon = (_eval_field(sp, psi, lat, lon + h) - _eval_field(sp, psi, lat, lon - h)) / (2 * h)
    dchi_dlon = (_eval_field(sp, chi, lat, lon + h) - _eval_field(sp, chi, lat, lon - h)) / (2 * h)
    U_ref = -np.cos(lat) * dpsi_dlat + dchi_dlon
    V_ref = dpsi_dlon + np.cos(lat) * dchi_dlat
    np.testing.assert_allclose(_eval_field(sp, U[0], lat, lon), U_ref, atol=1e-8)
    np.testing.assert_allclose(_eval_field(sp, V[0], lat, lon), V_ref, atol=1e-8)


@pytest.mark.parametrize("kind, n, lmax, mmax", [("octahedral", 8, 12, 11), ("full", 6, 9, 9), ("octahedral", 12, 17, 16)])
@pytest.mark.parametrize("unscale", [False, True])
def test_pullbacks_are_adjoints(kind, n, lmax, mmax, unscale):
    """<g_bar, S a> = Re<S^T g_bar, a> and Re<a_bar, A f> = <A^T a_bar, f> for random a, f,
    g_bar, a_bar: the pullbacks are the exact adjoints of synthesis S and analysis A."""
    sp = Spectrum(lmax, mmax)
    o = OracleTransform(sp, Grid(kind, n))
    rng = np.random.default_rng(7)
    K = 3
    a = rng.standard_normal((K, sp.nharm)) + 1j * rng.standard_normal((K, sp.nharm))
    gbar = rng.standard_normal((K, o.grid.npoints))
    lhs = np.sum(gbar * o.synthesis(a, unscale_coslat=unscale))
    rhs = np.real(np.sum(np.conj(o.spec2grid_pullback(gbar, unscale_coslat=unscale)) * a))
    assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(lhs))
    if not unscale:
        f = rng.standard_normal((K, o.grid.npoints))
        abar = rng.standard_normal((K, sp.nharm)) + 1j * rng.standard_normal((K, sp.nharm))
        lhs = np.real(np.sum(np.conj(abar) * o.analysis(f)))
        rhs = np.sum(o.grid2spec_pullback(abar) * f)
        assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(lhs))
    # a planted mistake (scale 1 on the interior bins) breaks the identity
    bad = o.forward_legendre(*[x / o.solid_angles[:n, None, None] for x in o.forward_fourier(gbar)])
    assert abs(np.sum(gbar * o.synthesis(a)) - np.real(np.sum(np.conj(bad) * a))) > 1e-6



def _eval_order(sp, a, m, phi):
    """sum_l a_lm lambda_lm(sin phi) for one order m at latitudes phi: (K, len(phi))."""
    lam = legendre_lambda(sp.lmax, sp.mmax, np.sin(phi))
    o, L = sp.offsets[m], sp.lmax - m
    return np.atleast_2d(a)[:, o:o + L] @ lam[o:o + L]


def _d1(F, phi, h):
    return (-F(phi + 2 * h) + 8 * F(phi + h) - 8 * F(phi - h) + F(phi - 2 * h)) / (12 * h)


def _d2(F, phi, h):
    return (-F(phi + 2 * h) + 16 * F(phi + h) - 30 * F(phi) + 16 * F(phi - h) - F(phi - 2 * h)) / (12 * h * h)


@pytest.mark.parametrize("lmax, mmax", [(12, 12), (13, 9)])
def test_gradient_operators_against_finite_differences(lmax, mmax):
    """The spectral operators reproduce the continuous ones on the fields they act on,
    evaluated at latitudes off the Gaussian grid with 4th-order differences in latitude:
    dpdx = dp/dlambda, dpdy = cos(phi) dp/dphi, Laplacian, divergence and curl of
    (u, v) given as (u, v) / cos(phi).  Inputs have zero top degrees, so the truncation to
    lmax degrees drops nothing."""
    sp = Spectrum(lmax, mmax)
    rng = np.random.default_rng(3)
    top = sp.l_of >= lmax - 2

    def rand():
        a = rng.standard_normal((2, sp.nharm)) + 1j * rng.standard_normal((2, sp.nharm))
        a[:, top] = 0.0
        return a

    p, u, v = rand(), rand(), rand()
    phi, h = np.linspace(-1.2, 1.2, 7), 1e-3
    c, sn = np.cos(phi), np.sin(phi)
    dpdx, dpdy = o_gradient(sp, p)
    lap = o_laplace(sp, p)
    div, cur = o_divergence(sp, u, v), o_divergence(sp, u, v, curl=True)
    for m in range(mmax):
        P = lambda x: _eval_order(sp, p, m, x)
        U = lambda x: _eval_order(sp, u, m, x)
        V = lambda x: _eval_order(sp, v, m, x)
        checks = [
            (_eval_order(sp, dpdx, m, phi), 1j * m * P(phi)),
            (_eval_order(sp, dpdy, m, phi), c * _d1(P, phi, h)),
            (_eval_order(sp, lap, m, phi), _d2(P, phi, h) - sn / c * _d1(P, phi, h) - m * m / c ** 2 * P(phi)),
            # div = dU/dlambda + (1/cos) d(cos^2 V)/dphi, curl = dV/dlambda - (1/cos) d(cos^2 U)/dphi
            (_eval_order(sp, div, m, phi), 1j * m * U(phi) - 2 * sn * V(phi) + c * _d1(V, phi, h)),
            (_eval_order(sp, cur, m, phi), 1j * m * V(phi) + 2 * sn * U(phi) - c * _d1(U, phi, h)),
        ]
        for got, want in checks:
            assert np.abs(got - want).max() < 1e-6 * max(1.0, np.abs(want).max())
    # inverse Laplacian, flipsign, add and radius as the reference orders them
    a = rng.standard_normal((2, sp.nharm)) + 1j * rng.standard_normal((2, sp.nharm))
    back = o_laplace(sp, o_laplace(sp, a), inverse=True)
    assert np.allclose(back[:, sp.l_of > 0], a[:, sp.l_of > 0], rtol=1e-13, atol=1e-13)
    assert np.all(back[:, sp.l_of == 0] == 0)
    assert np.allclose(o_laplace(sp, a, flipsign=True, add=a, radius=2.0), (a - o_laplace(sp, a)) / 4)
    d0 = o_divergence(sp, u, v)
    assert np.allclose(o_divergence(sp, u, v, flipsign=True, add=a, radius=3.0)[:, :-1], (a - d0)[:, :-1] / 3)
    # a planted mistake (the curl's signs in the divergence) fails the check above
    bad = 1j * gradient_factors(sp)[4] * u - (-gradient_factors(sp)[0] * np.roll(v, 1, axis=1))
    assert np.abs(_eval_order(sp, bad, 1, phi) - _eval_order(sp, div, 1, phi)).max() > 1e-3


@pytest.mark.parametrize("kind, n, lmax, mmax, shortcut", [
    ("full", 24, 33, 32, "lincubcoslat"),       # not monotonic in the ring index
    ("octahedral", 24, 33, 32, "linquadcoslat2"),
    ("octahedral", 48, 65, 64, "lincubcoslat"),
    ("full", 24, 33, 32, "quadratic"),
])
def test_shortcut_keeps_each_rings_own_orders(kind, n, lmax, mmax, shortcut):
    """Every ring keeps exactly its orders 0..mt_j (legendre_shortcuts.jl:21-49,
    spectral_transform.jl:154-155), pinned ring by ring against the linear-shortcut transform
    (which keeps every order the ring holds): synthesis on ring j equals the linear synthesis
    of the coefficients with orders above mt_j zeroed, and analysis equals the sum over rings
    of the linear analysis of the field on that ring pair alone, orders above mt_j zeroed."""
    sp, grid = Spectrum(lmax, mmax), Grid(kind, n)
    o = OracleTransform(sp, grid, shortcut)
    lin = OracleTransform(sp, grid, "linear")
    mt = o.mtrunc
    assert (mt <= lin.mtrunc).all()
    if shortcut == "lincubcoslat" and kind == "full":
        assert (np.diff(mt) < 0).any()  # the case the ring-set bookkeeping has to cope with
    # between the linear shortcut (cos lat -> 0) and the quadratic / cubic one (cos lat -> 1)
    lo = {"linquadcoslat2": 3, "lincubcoslat": 4, "quadratic": 3}[shortcut]
    assert (mt >= np.minimum(grid.nlons_half // lo, mmax - 1)).all()
    rng = np.random.default_rng(3)
    K = 3
    a = rng.standard_normal((K, sp.nharm)) + 1j * rng.standard_normal((K, sp.nharm))
    a[:, sp.offsets[0]:sp.offsets[0] + lmax] = a[:, sp.offsets[0]:sp.offsets[0] + lmax].real
    f = o.synthesis(a)
    m_of = np.repeat(np.arange(mmax), [lmax - m for m in range(mmax)])
    f_ref = np.zeros_like(f)
    for j in range(n):
        fj = lin.synthesis(np.where(m_of[None, :] <= mt[j], a, 0))
        for r in (j, grid.nlat - 1 - j):
            s0 = grid.ring_start[r]
            f_ref[:, s0:s0 + grid.nlons[r]] = fj[:, s0:s0 + grid.nlons[r]]
    assert np.allclose(f, f_ref, rtol=0, atol=1e-12 * np.abs(f_ref).max())
    g = rng.standard_normal((K, grid.npoints))
    b = o.analysis(g)
    b_ref = np.zeros_like(b)
    for j in range(n):
        gj = np.zeros_like(g)
        for r in (j, grid.nlat - 1 - j):
            s0 = grid.ring_start[r]
            gj[:, s0:s0 + grid.nlons[r]] = g[:, s0:s0 + grid.nlons[r]]
        b_ref += np.where(m_of[None, :] <= mt[j], lin.analysis(gj), 0)
    assert np.allclose(b, b_ref, rtol=0, atol=1e-12 * np.abs(b_ref).max())
    # the C restatement keeps the same orders
    c = CRefTransform(o)
    assert np.allclose(c.synthesis(a, np.float64), f, rtol=0, atol=1e-11 * np.abs(f).max())
    assert np.allclose(c.analysis(g, np.float64), b, rtol=0, atol=1e-11 * np.abs(b).max())


def test_shortcut_planted_mistake_fails():
    """The ring-by-ring check above notices a ring that keeps one order too many."""
    sp, grid = Spectrum(33, 32), Grid("full", 24)
    o = OracleTransform(sp, grid, "lincubcoslat")
    lin = OracleTransform(sp, grid, "linear")
    o.mtrunc = o.mtrunc.copy()
    j = int(np.argmin(o.mtrunc))
    o.mtrunc[j] += 1
    a = np.zeros((1, sp.nharm), complex)
    m = o.mtrunc[j]
    a[0, sp.offsets[m]] = 1.0  # l = m: nonzero on every ring
    f = o.synthesis(a)
    s0 = grid.ring_start[j]
    m_of = np.repeat(np.arange(32), [33 - mm for mm in range(32)])
    fj = lin.synthesis(np.where(m_of[None, :] <= m - 1, a, 0))
    assert not np.allclose(f[:, s0:s0 + grid.nlons[j]], fj[:, s0:s0 + grid.nlons[j]])
