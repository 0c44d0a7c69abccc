"""Minimal dry primitive-equation dynamical core on the MI355X transform engine.

A dynamics-only restatement of the reference's PrimitiveDryModel time step, used to
generate training data for the learned column correction and to run it: no physics, no
orography (NoOrography), dry, sigma coordinates with equally spaced levels, explicit
leapfrog with the Robert-Asselin-Williams filter and implicit hyperdiffusion.

Reference formulation (SpeedyWeather.jl):
  tendency pipeline          SpeedyWeather/src/dynamics/tendencies.jl:119-177
  pressure gradient flux     tendencies.jl:182-213
  vertical integration,
  vertical velocity          tendencies.jl:231-473
  vertical advection         SpeedyWeather/src/dynamics/vertical_advection.jl:23-185 (centred)
  geopotential               SpeedyWeather/src/dynamics/geopotential.jl:42-209
  vorticity / divergence
  flux form                  tendencies.jl:513-768, 897-998, 1123-1212
  mass conservation          SpeedyWeather/src/dynamics/horizontal_diffusion.jl:345 (l = 0 of ln ps)
  hyperdiffusion             horizontal_diffusion.jl:29-250 (power 4, 4 h / 1 h at T32)
  leapfrog + RAW filter      SpeedyWeather/src/time_stepping/steppers/leapfrog.jl:175-286
  sigma levels               SpeedyWeather/src/dynamics/vertical_coordinates.jl:83-87
  constants                  SpeedyWeather/src/dynamics/atmosphere.jl:14-74, planet.jl:4-28
  spectral truncation        LowerTriangularArrays/src/truncation.jl:67-89
The reference integrates the linear gravity-wave terms semi-implicitly
(implicit/implicit_primitive_equations.jl:69-352); this restatement is fully explicit with
a correspondingly shorter time step.

Spectral operators follow SpeedyTransforms/src/spectral_gradients.jl:30-125 (curl and
divergence of a vector whose grid components carry a 1/cos(lat) factor), :386-421
(Laplacian) and :483-527 (gradient), with the factors of gradient_arrays.jl:4-145.
Every spectral <-> grid transform runs on the engine; grid-point arithmetic uses torch.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from ._lib import lib
from .grids import OctahedralGaussianGrid, get_nlat_half
from .spectrum import LowerTriangularArray, Spectrum
from .transform import SpectralTransform, UV_from_vordiv_, transform_

RADIUS = 6.371e6        # m
ROTATION = 7.29e-5      # 1/s
GRAVITY = 9.81          # m/s^2
R_DRY = 287.05          # J/kg/K
CP = 1004.0             # J/kg/K
KAPPA = R_DRY / CP


def _eps(l, m):
    l = np.asarray(l, np.float64)
    m = np.asarray(m, np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        e = np.sqrt(np.maximum(l * l - m * m, 0.0) / (4.0 * l * l - 1.0))
    return np.where(l > 0, e, 0.0)


class SpectralOps:
    """Linear operators on packed coefficients, tensors (K, nharm) complex64."""

    def __init__(self, spectrum: Spectrum, device):
        l, m = spectrum.degrees_orders()  # 0-based
        last = l == spectrum.lmax - 1
        n = spectrum.nharm
        idx = np.arange(n)
        self.prev = torch.as_tensor(np.maximum(idx - 1, 0), device=device)
        self.next = torch.as_tensor(np.minimum(idx + 1, n - 1), device=device)
        e_l, e_l1 = _eps(l, m), _eps(l + 1, m)
        f = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=device)
        # cos(lat) d/dlat of a scalar: -(l-1) eps_l q_{l-1} + (l+2) eps_{l+1} q_{l+1}
        self.dy1 = f(-(l - 1) * e_l)
        self.dy2 = f(np.where(last, 0.0, (l + 2) * e_l1))
        # divergence / curl of (u, v) / cos(lat): (l+1) eps_l and l eps_{l+1}, last row 0
        self.vd1 = f(np.where(last, 0.0, (l + 1) * e_l))
        self.vd2 = f(np.where(last, 0.0, l * e_l1))
        self.m = f(m)
        self.keep = f(np.where(last, 0.0, 1.0))
        self.eig = f(-l * (l + 1.0))
        self.l = torch.as_tensor(l, device=device)

    def dlambda(self, q):
        """d/dlon: i m q."""
        return torch.complex(-self.m * q.imag, self.m * q.real)

    def coslat_dlat(self, q):
        return self.dy1 * q[:, self.prev] + self.dy2 * q[:, self.next]

    def divergence(self, u, v):
        """div of the vector whose grid components, times 1/cos(lat), have coefficients u, v."""
        return (self.dlambda(u) - self.vd1 * v[:, self.prev] + self.vd2 * v[:, self.next]) * self.keep

    def curl(self, u, v):
        return (self.dlambda(v) + self.vd1 * u[:, self.prev] - self.vd2 * u[:, self.next]) * self.keep

    def laplace(self, q):
        return self.eig * q


def truncate(x: torch.Tensor, src: Spectrum, dst: Spectrum) -> torch.Tensor:
    """Coefficients of the smaller spectrum dst copied from src (truncation.jl:67-89)."""
    ls, ms = dst.degrees_orders()
    idx = torch.as_tensor([src.offset(int(m)) + int(l) - int(m) for l, m in zip(ls, ms)], device=x.device)
    return x[:, idx].contiguous()


def grid_latitudes(grid) -> np.ndarray:
    """Latitude (radians) of every grid point, rings north to south."""
    nlat = 2 * grid.nlat_half
    x = np.zeros(nlat)
    w = np.zeros(nlat)
    lib().sht_gauss_legendre(nlat, x.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    lat_ring = np.arcsin(x[::-1])  # ascending nodes -> north to south
    return np.repeat(lat_ring, grid.nlons)


class DryPrimitiveModel:
    """Dry primitive equations at spectral truncation `trunc` (0-based, e.g. 31 or 127) on
    the octahedral Gaussian grid with quadratic dealiasing, `nlayers` sigma levels, for
    `nmembers` independent states side by side (columns of every transform)."""

    def __init__(self, trunc: int, nlayers: int = 8, nmembers: int = 1, dt: float | None = None,
                 precision: str = "fp32", device="cuda:0", radius: float = RADIUS):
        self.device = torch.device(device)
        self.trunc, self.L, self.E = trunc, nlayers, nmembers
        self.spectrum = Spectrum(trunc + 2, trunc + 1)  # one more degree for the vector path
        self.grid = OctahedralGaussianGrid(get_nlat_half(trunc + 1))
        self.a = radius
        L, E = nlayers, nmembers
        self.S = SpectralTransform(self.spectrum, self.grid, nlayers=E * (6 * L + 1), precision=precision,
                                   device=self.device.index or 0)
        self.ops = SpectralOps(self.spectrum, self.device)
        lat = grid_latitudes(self.grid)
        self.coslat = torch.as_tensor(np.cos(lat), dtype=torch.float32, device=self.device)
        self.coslat_inv = 1.0 / self.coslat
        self.f = torch.as_tensor(2 * ROTATION * np.sin(lat), dtype=torch.float32, device=self.device)
        self.lat = lat
        # sigma: half levels k/L, full levels (k - 1/2)/L, top to bottom
        sh = np.arange(L + 1) / L
        sf = (np.arange(L) + 0.5) / L
        self.dsigma = 1.0 / L
        self.sigma_half = torch.as_tensor(sh, dtype=torch.float32, device=self.device)
        self.sigma_full = torch.as_tensor(sf, dtype=torch.float32, device=self.device)
        # hydrostatic geopotential, linear in T: Phi_k = sum_j G[k, j] T_j (no orography)
        G = np.zeros((L, L))
        for k in range(L):
            G[k, k] = R_DRY * math.log(sh[k + 1] / sf[k])
            for j in range(k + 1, L):
                G[k, j] = R_DRY * math.log(sh[j + 1] / sh[j])
        self.G = torch.as_tensor(G, dtype=torch.float32, device=self.device)
        # time step: explicit leapfrog, gravity-wave limited
        self.dt = float(dt) if dt else 300.0 * 32.0 / (trunc + 1)
        # hyperdiffusion (l(l+1)/(T(T+1)))^4 with 4 h (1 h for divergence) at T32, scaled
        T1 = trunc + 1
        scale = 32.0 / T1
        ll = self.ops.l.to(torch.float64)
        shape = (ll * (ll + 1) / (trunc * (trunc + 1.0))) ** 4
        self.nu = (shape / (4 * 3600.0 * scale)).to(torch.float32)
        self.nu_div = (shape / (3600.0 * scale)).to(torch.float32)
        self.robert, self.williams = 0.1, 0.53
        self.state_prev = None
        self.state = None
        self.nsteps = 0

    # ------------------------------------------------------------------ transforms
    def _lta(self, t):
        return LowerTriangularArray(t.contiguous(), self.spectrum)

    def synth(self, coeffs: torch.Tensor, unscale=False) -> torch.Tensor:
        from .grids import Field
        out = torch.empty((coeffs.shape[0], self.grid.npoints), device=self.device)
        transform_(Field(out, self.grid), self._lta(coeffs), self.S, unscale_coslat=unscale)
        return out

    def analyse(self, field: torch.Tensor) -> torch.Tensor:
        from .grids import Field
        out = torch.empty((field.shape[0], self.spectrum.nharm), dtype=torch.complex64, device=self.device)
        transform_(self._lta(out), Field(field.contiguous(), self.grid), self.S)
        return out

    def uv(self, vor: torch.Tensor, div: torch.Tensor):
        """Spectral u cos(lat), v cos(lat) from vorticity and divergence (radius scaled)."""
        U = torch.empty_like(vor)
        V = torch.empty_like(div)
        UV_from_vordiv_(self._lta(U), self._lta(V), self._lta(vor), self._lta(div), self.S, radius=self.a)
        return U, V

    # ------------------------------------------------------------------ dynamics
    def tendencies(self, st: dict, with_grid: bool = False):
        """Spectral tendencies of vor, div, temp (E*L, nharm) and lnps (E, nharm)."""
        L, E, a = self.L, self.E, self.a
        ops = self.ops
        vor, div, temp, lnps = st["vor"], st["div"], st["temp"], st["lnps"]
        U, V = self.uv(vor, div)
        g1 = self.synth(torch.cat([U, V, ops.dlambda(lnps), ops.coslat_dlat(lnps)]), unscale=True)
        g2 = self.synth(torch.cat([temp, div, vor, lnps]))
        EL = E * L
        u, v = g1[:EL].view(E, L, -1), g1[EL:2 * EL].view(E, L, -1)
        dqdx, dqdy = (g1[2 * EL:2 * EL + E] / a).view(E, 1, -1), (g1[2 * EL + E:] / a).view(E, 1, -1)
        T, D, Z = g2[:EL].view(E, L, -1), g2[EL:2 * EL].view(E, L, -1), g2[2 * EL:3 * EL].view(E, L, -1)
        q = g2[3 * EL:].view(E, 1, -1)
        ds = self.dsigma
        uq = u * dqdx + v * dqdy                      # (u, v) . grad ln ps
        dsum = ds * (D + uq)                          # layer mass-flux divergence
        csum = torch.cumsum(dsum, dim=1)
        lnps_t = -csum[:, -1:, :]                     # d ln ps / dt
        sh = self.sigma_half.view(1, -1, 1)
        sdot = torch.zeros((E, L + 1, u.shape[-1]), device=self.device)
        sdot[:, 1:L] = -csum[:, :L - 1] - sh[:, 1:L] * lnps_t
        def vadv(X):
            dX = X[:, 1:] - X[:, :-1]                 # (E, L-1)
            out = torch.zeros_like(X)
            out[:, :-1] += sdot[:, 1:L] * dX
            out[:, 1:] += sdot[:, 1:L] * dX
            return out / (2 * ds)
        sf = self.sigma_full.view(1, -1, 1)
        omp = uq - (csum - 0.5 * dsum) / sf           # omega / p on full levels
        absvor = Z + self.f
        Au = absvor * v - vadv(u) - R_DRY * T * dqdx
        Av = -absvor * u - vadv(v) - R_DRY * T * dqdy
        ke = 0.5 * (u * u + v * v)
        Tt = T * D - vadv(T) + KAPPA * T * omp
        ci = self.coslat_inv
        flds = torch.cat([(Au * ci).view(EL, -1), (Av * ci).view(EL, -1), (u * T * ci).view(EL, -1),
                          (v * T * ci).view(EL, -1), ke.view(EL, -1), Tt.view(EL, -1), lnps_t.view(E, -1)])
        s = self.analyse(flds)
        Aus, Avs, uTs, vTs, kes, Tts = (s[i * EL:(i + 1) * EL] for i in range(6))
        lnps_ts = s[6 * EL:].clone()
        phi = torch.einsum("kj,ejn->ekn", self.G.to(torch.complex64),
                           temp.view(E, L, -1)).reshape(EL, -1)
        vor_t = ops.curl(Aus, Avs) / a
        div_t = ops.divergence(Aus, Avs) / a - ops.laplace(phi + kes) / (a * a)
        temp_t = -ops.divergence(uTs, vTs) / a + Tts
        lnps_ts[:, 0] = 0                              # global mass conservation
        # the extra degree only serves the vector path: no prognostic tendency there
        keep = ops.keep
        tend = {"vor": vor_t * keep, "div": div_t * keep, "temp": temp_t * keep, "lnps": lnps_ts * keep}
        if with_grid:
            return tend, {"u": u, "v": v, "T": T, "lnps": q}
        return tend

    # ------------------------------------------------------------------ time stepping
    def _diffuse(self, name, x, dt2):
        nu = self.nu_div if name == "div" else self.nu
        if name == "lnps":
            return x
        return x / (1.0 + dt2 * nu)

    def step(self, correction=None):
        """One leapfrog step (Euler forward for the first); `correction(model, tend, grid)`
        may add to the spectral tendencies."""
        need_grid = correction is not None
        out = self.tendencies(self.state, with_grid=need_grid)
        tend, grid = (out if need_grid else (out, None))
        if correction is not None:
            correction(self, tend, grid)
        if self.state_prev is None:
            dt2, base = self.dt, self.state
        else:
            dt2, base = 2.0 * self.dt, self.state_prev
        new = {k: self._diffuse(k, base[k] + dt2 * tend[k], dt2) for k in self.state}
        if self.state_prev is not None:
            w1 = self.robert * self.williams / 2
            w2 = self.robert * (1 - self.williams) / 2
            for k in self.state:
                d = self.state_prev[k] - 2 * self.state[k] + new[k]
                self.state[k] = self.state[k] + w1 * d
                new[k] = new[k] - w2 * d
        self.state_prev, self.state = self.state, new
        self.nsteps += 1

    # ------------------------------------------------------------------ initial conditions
    def jet_state(self, seeds, u0=35.0, lat0=30.0, width=10.0, T0=288.0, lapse=0.0065,
                  noise_K=1.0, noise_trunc=None) -> dict:
        """Zonal jets at +-lat0 (peak u0 m/s, Gaussian in latitude) over an atmosphere at
        rest in the vertical mean, T0 at the surface with a constant lapse rate, plus
        N(0, noise_K) temperature noise per grid point from each member's seed.  With
        noise_trunc, the noise keeps only degrees l <= noise_trunc and is rescaled to
        noise_K root mean square per level."""
        L, E = self.L, self.E
        assert len(seeds) == E
        latd = np.degrees(self.lat)
        ujet = u0 * (np.exp(-((latd - lat0) / width) ** 2) + np.exp(-((latd + lat0) / width) ** 2))
        sig = self.sigma_full.cpu().numpy()
        u = np.tile(ujet, (E * L, 1)).astype(np.float32)
        Tprof = T0 * sig ** (R_DRY * lapse / GRAVITY)
        T = np.empty((E * L, self.grid.npoints), np.float32)
        for e, seed in enumerate(seeds):
            rng = np.random.default_rng(int(seed))
            T[e * L:(e + 1) * L] = Tprof[:, None] + noise_K * rng.standard_normal((L, self.grid.npoints))
        u_t = torch.as_tensor(u, device=self.device)
        us = self.analyse(u_t * self.coslat_inv)
        zero = torch.zeros_like(us)
        vor = self.ops.curl(us, zero) / self.a
        div = torch.zeros_like(vor)
        if noise_trunc is not None:
            Tn = torch.as_tensor(T - Tprof.astype(np.float32)[None, :, None].repeat(E, 0).reshape(E * L, 1),
                                 device=self.device)
            ns = self.analyse(Tn) * (self.ops.l <= noise_trunc).to(torch.float32)
            Tn = self.synth(ns)
            Tn = Tn * (noise_K / Tn.pow(2).mean(dim=1, keepdim=True).sqrt().clamp_min(1e-30))
            T = Tn.cpu().numpy() + Tprof.astype(np.float32)[None, :, None].repeat(E, 0).reshape(E * L, 1)
        temp = self.analyse(torch.as_tensor(T, device=self.device))
        lnps = self.analyse(torch.full((E, self.grid.npoints), math.log(1e5), device=self.device))
        keep = self.ops.keep
        return {"vor": vor * keep, "div": div * keep, "temp": temp * keep, "lnps": lnps * keep}

    def set_state(self, st: dict):
        self.state = {k: v.clone() for k, v in st.items()}
        self.state_prev = None
        self.nsteps = 0

    def run(self, nsteps: int, correction=None):
        for _ in range(nsteps):
            self.step(correction)

    def close(self):
        self.S.close()
