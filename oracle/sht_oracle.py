"""FP64 CPU restatement of SpeedyTransforms' spherical harmonic transform.

TEST INFRASTRUCTURE ONLY. This module is the CPU reference the HIP engine is compared
against; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it. The product (``speedyweather.jl_amd``) never imports or
calls anything under ``oracle/``.

It restates, loop structure for loop structure, the reference CPU path
(all citations relative to the SpeedyWeather.jl repository root):

* sizes: ``SpeedyTransforms/src/aliasing.jl:9-59`` (get_nlat_half, roundup_fft),
  ``RingGrids/src/grids/octahedral_gaussian.jl:50-116`` (16+4j points per ring,
  north-to-south ring order, mirrored south), ``RingGrids/src/grids/full_gaussian.jl:43-61``;
* spectral packing: ``LowerTriangularArrays/src/lower_triangular_array.jl:367-374``
  (m-major columns, l within a column);
* Gaussian latitudes and weights: ``RingGrids/src/grids/full_gaussian.jl:48-50``,
  ``RingGrids/src/quadrature_weights.jl:11,48-49`` (solid angles w_j * 2pi/nlon_j);
* Legendre polynomials: ``SpeedyTransforms/src/spectral_transform.jl:167-172`` calling
  ``Legendre.λlm!`` of AssociatedLegendrePolynomials.jl v1.0.2 (not vendored): the
  spherical-harmonic normalised P_l^m with the Condon-Shortley phase, restated here by
  the textbook three-term recursion in FP64;
* Legendre shortcuts: ``SpeedyTransforms/src/legendre_shortcuts.jl:21-49`` (all five);
* inverse Legendre: ``SpeedyTransforms/src/legendre.jl:2-33,38-101``;
* forward Legendre: ``SpeedyTransforms/src/legendre.jl:104-126,131-198``;
* Fourier stage: ``SpeedyTransforms/src/fourier.jl:27-77,129-250`` with FFTW's
  unnormalised rfft / brfft (FFTW.jl 1.10.0, not vendored), restated with numpy.fft
  (rfft; irfft*N, which like brfft ignores Im of the DC and Nyquist bins);
* vector path: ``SpeedyTransforms/src/spectral_gradients.jl:309-366`` with the factors
  of ``SpeedyTransforms/src/gradient_arrays.jl:4-21,98-145``;
* adjoints (pullbacks): ``SpeedyTransforms/ext/SpeedyTransformsEnzymeExt.jl:33-59,263-319``;
* spectral gradient operators: ``SpeedyTransforms/src/spectral_gradients.jl:30-125,385-527``
  with ``SpeedyTransforms/src/gradient_arrays.jl:4-100,150-170``.

Pinning (tests/test_oracle.py): Gauss nodes/weights against numpy's leggauss and exact
polynomial quadrature; Λ against closed forms for l <= 2, against scipy's spherical
harmonics, discrete orthonormality and the recursion residual; the FFT edge semantics
against a brute-force DFT; the whole transform against Y00 -> constant grid, single
harmonic round trips, the FP64 grid->spec->grid round trip at cubic dealiasing, the
vector path against solid-body rotation, the pullbacks against the dot-product identity
with synthesis and analysis (they are their exact adjoints), the gradient operators against
finite differences in latitude of the fields they act on, evaluated at arbitrary latitudes.
The Condon-Shortley phase: the reference's documentation states the real-field symmetry
a_{l,-m} = (-1)^m conj(a_{l,m}) (docs/src/spectral_transform.md:66-69), which holds for
the phase-carrying convention that scipy's sph_harm_y also uses (the pin above); no
reference test fixes the sign of odd-m polynomials numerically, and round trips do not
depend on it.
"""
from __future__ import annotations

import math

import numpy as np

# ----------------------------------------------------------------------------- sizes


def roundup_fft(n: int, small_primes=(2, 3, 5)) -> int:
    """Smallest even integer >= n whose prime factors are all in small_primes
    (SpeedyTransforms/src/aliasing.jl:42-59)."""
    n += n % 2
    while True:
        r = n
        for p in small_primes:
            while r % p == 0:
                r //= p
        if r == 1:
            return n
        n += 2


def get_nlat_half(truncation: int, dealiasing: float = 2.0) -> int:
    """aliasing.jl:9-15, truncation 1-based (= mmax of the Spectrum)."""
    return roundup_fft(math.ceil(((1 + dealiasing) * (truncation - 1) + 1) / 4))


def get_truncation(nlat_half: int, dealiasing: float = 2.0) -> int:
    """aliasing.jl:23-29."""
    return math.floor((4 * nlat_half - 1) / (dealiasing + 1)) + 1


class Spectrum:
    """1-based (lmax, mmax) as LowerTriangularArrays/src/spectrum.jl:6-70."""

    def __init__(self, lmax: int, mmax: int):
        assert lmax >= mmax >= 1
        self.lmax, self.mmax = lmax, mmax
        self.nharm = lmax * mmax - mmax * (mmax - 1) // 2
        m = np.arange(mmax)
        self.offsets = m * lmax - m * (m - 1) // 2  # 0-based start of column m
        ls, ms = [], []
        for mm in range(mmax):
            for ll in range(mm, lmax):
                ls.append(ll)
                ms.append(mm)
        self.l_of = np.array(ls)
        self.m_of = np.array(ms)

    def index(self, l: int, m: int) -> int:
        """0-based packed index of degree l, order m (lower_triangular_array.jl:367)."""
        return int(self.offsets[m] + (l - m))


class Grid:
    """Octahedral or full Gaussian ring grid (octahedral_gaussian.jl:50-116,
    full_gaussian.jl:43-61): nlat = 2 nlat_half rings north to south, south mirrored."""

    def __init__(self, kind: str, nlat_half: int):
        assert kind in ("octahedral", "full")
        self.kind, self.nlat_half = kind, nlat_half
        n = nlat_half
        if kind == "octahedral":
            north = [16 + 4 * j for j in range(1, n + 1)]
        else:
            north = [4 * n] * n
        self.nlons_half = np.array(north)
        self.nlons = np.array(north + north[::-1])
        self.ring_start = np.concatenate([[0], np.cumsum(self.nlons)[:-1]])
        self.npoints = int(self.nlons.sum())
        self.nlat = 2 * n


# ------------------------------------------------------------------ quadrature, Λ


def _legendre_p_and_dp(N: int, x: np.ndarray):
    p0 = np.ones_like(x)
    p1 = x.copy()
    for k in range(2, N + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = N * (x * p1 - p0) / (x * x - 1)
    return p1, dp


def gauss_legendre(N: int):
    """Nodes (ascending) and weights of N-point Gauss-Legendre quadrature, Newton
    iteration on P_N in FP64 (FastGaussQuadrature.gausslegendre's contract,
    RingGrids/src/grids/full_gaussian.jl:49)."""
    i = np.arange(1, N + 1)
    x = -np.cos(np.pi * (i - 0.25) / (N + 0.5))
    for _ in range(100):
        p, dp = _legendre_p_and_dp(N, x)
        dx = p / dp
        x = x - dx
        if np.max(np.abs(dx)) < 1e-16:
            break
    p, dp = _legendre_p_and_dp(N, x)
    w = 2.0 / ((1 - x * x) * dp * dp)
    return x, w


def legendre_lambda(lmax: int, mmax: int, x) -> np.ndarray:
    """λ_l^m(x) for 0 <= m < mmax, m <= l < lmax, packed as a LowerTriangularArray
    column (m-major); x scalar or 1-D array (then one column per x). Orthonormal on the
    sphere (λ_0^0 = 1/(2 sqrt(pi))), Condon-Shortley phase (-1)^m. Three-term recursion
    in l from the diagonal λ_m^m, all in FP64."""
    xa = np.atleast_1d(np.asarray(x, dtype=np.float64))
    s = np.sqrt(np.maximum(0.0, 1.0 - xa * xa))
    out = np.zeros((lmax * mmax - mmax * (mmax - 1) // 2, xa.size))
    pmm = np.full(xa.size, 1.0 / math.sqrt(4 * math.pi))
    idx = 0
    for m in range(mmax):
        if m > 0:
            pmm = -math.sqrt((2 * m + 1) / (2.0 * m)) * s * pmm
        out[idx] = pmm
        if m + 1 < lmax:
            p0 = pmm
            p1 = math.sqrt(2 * m + 3) * xa * pmm
            out[idx + 1] = p1
            for l in range(m + 2, lmax):
                a = math.sqrt((4.0 * l * l - 1) / (l * l - m * m))
                b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1) ** 2 - 1))
                p0, p1 = p1, a * (xa * p1 - b * p0)
                out[idx + l - m] = p1
        idx += lmax - m
    return out[:, 0] if np.ndim(x) == 0 else out


SHORTCUTS = ("linear", "quadratic", "cubic", "linquadcoslat2", "lincubcoslat")


def legendre_shortcut(name: str, nlon, latd) -> np.ndarray:
    """Last order kept per ring (SpeedyTransforms/src/legendre_shortcuts.jl:21-49): nlon ÷ 2,
    ÷ 3, ÷ 4, floor(nlon / (2 + cosd(latd)^2)), floor(nlon / (2 + 2 cosd(latd)))."""
    nlon = np.asarray(nlon, dtype=np.int64)
    c = np.cos(np.radians(np.asarray(latd, dtype=np.float64)))
    if name == "linear":
        return nlon // 2
    if name == "quadratic":
        return nlon // 3
    if name == "cubic":
        return nlon // 4
    if name == "linquadcoslat2":
        return np.floor(nlon / (2.0 + c * c)).astype(np.int64)
    if name == "lincubcoslat":
        return np.floor(nlon / (2.0 + 2.0 * c)).astype(np.int64)
    raise ValueError(f"unknown Legendre shortcut {name}")


class OracleTransform:
    """FP64 restatement of SpectralTransform + transform! (spectral_transform.jl:118-302,
    515-631) for octahedral / full Gaussian grids with any of the five Legendre shortcuts
    (each ring keeps its own orders, also where the latitude-dependent ones are not
    monotonic in the ring index)."""

    def __init__(self, spectrum: Spectrum, grid: Grid, shortcut: str = "linear"):
        self.spectrum, self.grid = spectrum, grid
        n = grid.nlat_half
        x, w = gauss_legendre(2 * n)
        latd = np.degrees(np.arccos(x)) - 90.0  # full_gaussian.jl:49 (north first)
        self.latd = latd
        colat = np.radians(90.0 - latd)
        self.cos_colat = np.cos(colat)  # spectral_transform.jl:147-149
        self.coslat = np.cos(np.radians(latd))  # :150
        self.coslat_inv = 1.0 / self.coslat
        self.weights = w
        self.solid_angles = w * 2 * np.pi / grid.nlons  # quadrature_weights.jl:48-49
        self.mtrunc = np.minimum(legendre_shortcut(shortcut, grid.nlons_half, latd[:n]),
                                 spectrum.mmax - 1)  # spectral_transform.jl:154-155
        self.nfreq = grid.nlons_half // 2 + 1
        # Λ.data[lm, j] for the northern rings (spectral_transform.jl:167-172)
        self.legendre = legendre_lambda(spectrum.lmax, spectrum.mmax, self.cos_colat[:n])
        self.norm_sphere = 2 * math.sqrt(math.pi)

    # -- spectral -> grid (legendre.jl:38-101, fourier.jl:215-250)
    def inverse_legendre(self, coeffs: np.ndarray, unscale_coslat: bool = False):
        """coeffs (K, nharm) complex -> g_north, g_south (n, K, nfreq_max) complex."""
        sp, n = self.spectrum, self.grid.nlat_half
        K = coeffs.shape[0]
        nfmax = int(self.nfreq.max())
        gN = np.zeros((n, K, nfmax), complex)
        gS = np.zeros((n, K, nfmax), complex)
        for m in range(sp.mmax):
            o = sp.offsets[m]
            Lm = sp.lmax - m
            rings = np.nonzero(self.mtrunc >= m)[0]
            if rings.size == 0:
                continue
            lam = self.legendre[o:o + Lm][:, rings]  # (Lm, J)
            a = coeffs[:, o:o + Lm]  # (K, Lm)
            even = a[:, 0::2] @ lam[0::2]  # (K, J)
            odd = a[:, 1::2] @ lam[1::2]
            gN[rings, :, m] = (even + odd).T
            gS[rings, :, m] = (even - odd).T
        if unscale_coslat:
            gN *= self.coslat_inv[:n, None, None]
            gS *= self.coslat_inv[:n, None, None]
        return gN, gS

    def inverse_fourier(self, gN, gS):
        g = self.grid
        n = g.nlat_half
        K = gN.shape[1]
        field = np.zeros((K, g.npoints))
        for j in range(n):
            N = int(g.nlons_half[j])
            nf = N // 2 + 1
            js = g.nlat - 1 - j
            field[:, g.ring_start[j]:g.ring_start[j] + N] = np.fft.irfft(gN[j, :, :nf], n=N) * N
            field[:, g.ring_start[js]:g.ring_start[js] + N] = np.fft.irfft(gS[j, :, :nf], n=N) * N
        return field

    def synthesis(self, coeffs: np.ndarray, unscale_coslat: bool = False) -> np.ndarray:
        coeffs = np.atleast_2d(coeffs)
        return self.inverse_fourier(*self.inverse_legendre(coeffs, unscale_coslat))

    # -- grid -> spectral (fourier.jl:129-166, legendre.jl:131-198)
    def forward_fourier(self, field: np.ndarray):
        g = self.grid
        n = g.nlat_half
        K = field.shape[0]
        nfmax = int(self.nfreq.max())
        fN = np.zeros((n, K, nfmax), complex)
        fS = np.zeros((n, K, nfmax), complex)
        for j in range(n):
            N = int(g.nlons_half[j])
            nf = N // 2 + 1
            js = g.nlat - 1 - j
            fN[j, :, :nf] = np.fft.rfft(field[:, g.ring_start[j]:g.ring_start[j] + N])
            fS[j, :, :nf] = np.fft.rfft(field[:, g.ring_start[js]:g.ring_start[js] + N])
        return fN, fS

    def forward_legendre(self, fN, fS):
        sp, n = self.spectrum, self.grid.nlat_half
        K = fN.shape[1]
        coeffs = np.zeros((K, sp.nharm), complex)
        dO = self.solid_angles[:n]
        for m in range(sp.mmax):
            rings = np.nonzero(self.mtrunc >= m)[0]
            if rings.size == 0:
                continue
            o = sp.offsets[m]
            Lm = sp.lmax - m
            lam = self.legendre[o:o + Lm][:, rings]  # (Lm, J)
            even = dO[rings, None] * (fN[rings, :, m] + fS[rings, :, m])  # (J, K)
            odd = dO[rings, None] * (fN[rings, :, m] - fS[rings, :, m])
            out = np.zeros((Lm, K), complex)
            out[0::2] = lam[0::2] @ even
            out[1::2] = lam[1::2] @ odd
            coeffs[:, o:o + Lm] = out.T
        return coeffs

    def analysis(self, field: np.ndarray) -> np.ndarray:
        field = np.atleast_2d(field)
        return self.forward_legendre(*self.forward_fourier(field))

    # -- adjoints (SpeedyTransforms/ext/SpeedyTransformsEnzymeExt.jl:33-59,263-319)
    def adjoint_scale(self) -> np.ndarray:
        """(n, 1, nfreq_max): 2 for the interior rfft bins, 1 for DC and the Nyquist bin of
        even rings, 0 past a ring's bins (SpeedyTransformsEnzymeExt.jl:35-59)."""
        g, n = self.grid, self.grid.nlat_half
        s = np.zeros((n, 1, int(self.nfreq.max())))
        for j in range(n):
            N = int(g.nlons_half[j])
            nf = N // 2 + 1
            s[j, 0, :nf] = [2.0 if (0 < i < nf - 1 or (N % 2 == 1 and i > 0)) else 1.0 for i in range(nf)]
        return s

    def spec2grid_pullback(self, field_bar: np.ndarray, unscale_coslat: bool = False) -> np.ndarray:
        """Cotangent of the coefficients of synthesis (SpeedyTransformsEnzymeExt.jl:263-290):
        forward FFT, times the adjoint scale (and 1/cos(lat)), divided by the solid angle the
        forward Legendre applies, then the forward Legendre."""
        n = self.grid.nlat_half
        dgN, dgS = self.forward_fourier(np.atleast_2d(field_bar))
        w = self.adjoint_scale() / self.solid_angles[:n, None, None]
        if unscale_coslat:
            w = w * self.coslat_inv[:n, None, None]
        return self.forward_legendre(dgN * w, dgS * w)

    def grid2spec_pullback(self, coeffs_bar: np.ndarray) -> np.ndarray:
        """Cotangent of the grid of analysis (SpeedyTransformsEnzymeExt.jl:294-319): inverse
        Legendre, times the solid angle, divided by the adjoint scale, then the inverse FFT."""
        n = self.grid.nlat_half
        dfN, dfS = self.inverse_legendre(np.atleast_2d(coeffs_bar))
        s = self.adjoint_scale()
        w = np.where(s > 0, self.solid_angles[:n, None, None] / np.where(s > 0, s, 1.0), 0.0)
        return self.inverse_fourier(dfN * w, dfS * w)


# --------------------------------------------------------------- vector path (C2)


def uv_from_vordiv(spectrum: Spectrum, vor: np.ndarray, div: np.ndarray, radius: float = 1.0):
    """U, V = (u, v) cos(lat) spectra from vorticity and divergence
    (spectral_gradients.jl:309-366; factors gradient_arrays.jl:4-21,98-145).
    vor, div: (K, nharm) complex."""
    lmax, mmax = spectrum.lmax, spectrum.mmax

    def eps(l, m):
        return math.sqrt((l * l - m * m) / (4.0 * l * l - 1)) if l > 0 else 0.0

    vor = np.atleast_2d(vor)
    div = np.atleast_2d(div)
    U = np.zeros_like(vor)
    V = np.zeros_like(vor)
    for m in range(mmax):
        o = spectrum.offsets[m]
        for l in range(m, lmax):
            i = o + (l - m)
            x = -m / (l * (l + 1)) if l > 0 else 0.0
            u1 = eps(l, m) / l if l > 0 else 0.0
            u2 = eps(l + 1, m) / (l + 1) if l < lmax - 1 else 0.0
            z = 1j * x
            first = (i == 0)
            if first:
                U[:, i] = z * div[:, i] + u2 * vor[:, i + 1]
                V[:, i] = z * vor[:, i] - u2 * div[:, i + 1]
            elif l == lmax - 2:
                U[:, i] = z * div[:, i] - u1 * vor[:, i - 1]
                V[:, i] = z * vor[:, i] + u1 * div[:, i - 1]
            elif l == lmax - 1:
                U[:, i] = -u1 * vor[:, i - 1]
                V[:, i] = u1 * div[:, i - 1]
            else:
                up = vor[:, i + 1]
                dp = div[:, i + 1]
                vm = vor[:, i - 1] if l > m else 0.0
                dm = div[:, i - 1] if l > m else 0.0
                U[:, i] = z * div[:, i] + u2 * up - u1 * vm
                V[:, i] = z * vor[:, i] + u1 * dm - u2 * dp
    return U * radius, V * radius


# ------------------------------------------------------- spectral gradient operators


def gradient_factors(spectrum: Spectrum):
    """Per packed (l, m), FP64: grad_y_vordiv1, grad_y_vordiv2, grad_y1, grad_y2,
    grad_x_vordiv and m (SpeedyTransforms/src/gradient_arrays.jl:4-21,27-76,80-100): with
    0-based l, m and eps(l, m) = sqrt((l^2 - m^2) / (4 l^2 - 1)):
    gyvd1 = (l+1) eps(l, m), gyvd2 = l eps(l+1, m), gy1 = -(l-1) eps(l, m),
    gy2 = (l+2) eps(l+1, m), gxvd = m; every one but gy1 is 0 on the last degree l = lmax-1."""
    lmax, mmax = spectrum.lmax, spectrum.mmax

    def eps(l, m):
        return math.sqrt((l * l - m * m) / (4.0 * l * l - 1)) if l > 0 else 0.0

    f = np.zeros((6, spectrum.nharm))
    for m in range(mmax):
        for l in range(m, lmax):
            i = spectrum.offsets[m] + (l - m)
            top = l == lmax - 1
            f[0, i] = 0.0 if top else (l + 1) * eps(l, m)
            f[1, i] = 0.0 if top else l * eps(l + 1, m)
            f[2, i] = -(l - 1) * eps(l, m)
            f[3, i] = 0.0 if top else (l + 2) * eps(l + 1, m)
            f[4, i] = 0.0 if top else m
            f[5, i] = m
    return f


def _shift(a, d):
    """a[:, lm + d] over the packed index (zero past either end)."""
    out = np.zeros_like(a)
    if d > 0:
        out[:, :-d] = a[:, d:]
    else:
        out[:, -d:] = a[:, :d]
    return out


def divergence(spectrum: Spectrum, u, v, curl: bool = False, flipsign: bool = False, add=None, radius: float = 1.0):
    """∇⋅(u, v) or ∇×(u, v) of components carrying 1/cos(lat)
    (SpeedyTransforms/src/spectral_gradients.jl:30-125): the divergence kernel
    i m u - gyvd1 v[lm-1] + gyvd2 v[lm+1] (curl: on (v, u) with + gyvd1, - gyvd2), the last
    packed element 0, then flipsign, add and the 1/radius scaling in that order."""
    f = gradient_factors(spectrum)
    u, v = np.atleast_2d(u), np.atleast_2d(v)
    X, Y = (v, u) if curl else (u, v)
    sg = 1.0 if curl else -1.0
    val = 1j * f[4] * X + sg * (f[0] * _shift(Y, -1) - f[1] * _shift(Y, 1))
    val[:, -1] = 0.0
    if flipsign:
        val = -val
    if add is not None:
        val = val + np.atleast_2d(add)
        val[:, -1] = 0.0
    return val / radius


def laplace(spectrum: Spectrum, alms, inverse: bool = False, flipsign: bool = False, add=None, radius: float = 1.0):
    """∇² or ∇⁻² (the l = 0 mode set to 0) by the eigenvalues -l(l+1)
    (spectral_gradients.jl:385-420, gradient_arrays.jl:150-170), then flipsign, add, and the
    radius scaling (1/R^2, or R^2 for the inverse)."""
    l = spectrum.l_of.astype(np.float64)
    e = -l * (l + 1)
    if inverse:
        e = np.where(l == 0, 0.0, 1.0 / np.where(l == 0, 1.0, e))
    val = np.atleast_2d(alms) * e
    if flipsign:
        val = -val
    if add is not None:
        val = val + np.atleast_2d(add)
    return val * (radius ** 2 if inverse else 1.0 / radius ** 2)


def gradient(spectrum: Spectrum, p, radius: float = 1.0):
    """dpdx = i m p, dpdy = gy1 p[lm-1] + gy2 p[lm+1] (∇!, spectral_gradients.jl:483-527),
    both divided by the radius."""
    f = gradient_factors(spectrum)
    p = np.atleast_2d(p)
    return 1j * f[5] * p / radius, (f[2] * _shift(p, -1) + f[3] * _shift(p, 1)) / radius
