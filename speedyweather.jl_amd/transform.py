"""SpectralTransform and transform! on the MI355X engine.

Mirrors SpeedyTransforms' API (SpeedyTransforms/src/spectral_transform.jl):
  SpectralTransform(spectrum, grid; nlayers, LegendreShortcut)     :118-302
  transform!(field, coeffs, S; unscale_coslat)  ->  transform_(field, coeffs, S, ...)  :515-570
  transform!(coeffs, field, S)                  ->  transform_(coeffs, field, S)       :579-631
  transform(field, S) / transform(coeffs, S)    ->  transform(...)                     :638-682
  UV_from_vordiv!(U, V, vor, div, S; radius)    ->  UV_from_vordiv_(...)   spectral_gradients.jl:309-329
Every call goes through libsht's C ABI with device pointers on the caller's current
torch stream.  There is no CPU fallback: data must live on the plan's GPU.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import DimensionMismatch, check, lib
from .grids import Field, FullGaussianGrid, OctahedralGaussianGrid, get_nlat_half
from .spectrum import LowerTriangularArray, Spectrum


class SpectralTransform:
    """Precomputed transform between `spectrum` and `grid` for up to `nlayers` columns.

    precision: "fp32" (FP32 products and accumulation), "fp16x3" (scaled FP16 hi/lo split,
    three MFMA products, FP32 accumulation) or "bf16x3" (BF16 hi/lo split)."""

    def __init__(self, spectrum: Spectrum, grid, nlayers: int = 1, LegendreShortcut: str = "linear",
                 precision: str = "fp32", device: int | None = None, gpu_graphs: bool = False):
        if grid.kind not in _lib.GRIDS:
            raise NotImplementedError(f"{type(grid).__name__} is not supported by the MI355X engine")
        self.spectrum, self.grid = spectrum, grid
        self.nlayers = int(nlayers)
        self.LegendreShortcut = LegendreShortcut
        self.precision = precision
        if device is None:
            device = torch.cuda.current_device() if torch.cuda.is_available() else 0
        self.device = int(device)
        self._desc = _lib.ShtDesc(_lib.GRIDS[grid.kind], grid.nlat_half, spectrum.lmax, spectrum.mmax,
                                  self.nlayers, _lib.SHORTCUTS[LegendreShortcut],
                                  _lib.PRECISIONS[precision], self.device)
        self._plan = C.c_void_p()
        self._create()
        if gpu_graphs:
            self.set_graphs(True)

    def _create(self):
        check(lib().sht_plan_create(C.byref(self._desc), C.byref(self._plan)), "sht_plan_create")

    def set_graphs(self, enable: bool = True, force_failure: bool = False):
        """Graph replay of repeated calls (the reference's gpu_graphs, SpeedyTransforms/ext/
        gpu_graphs_common.jl:293-331): each distinct call's two stage kernels replay from a cached
        graph, at most 64 per direction (sht_plan_set_graphs)."""
        check(lib().sht_plan_set_graphs(self._plan, int(bool(enable)), 1 if force_failure else 0),
              "sht_plan_set_graphs")

    def graph_count(self) -> tuple:
        """(graphs cached, graphs that failed) (sht_plan_graph_count)."""
        n, f = C.c_int32(), C.c_int32()
        check(lib().sht_plan_graph_count(self._plan, C.byref(n), C.byref(f)), "sht_plan_graph_count")
        return n.value, f.value

    @classmethod
    def from_grid(cls, grid, truncation: int = 0, dealiasing: float = 2.0, one_more_degree: bool = False,
                  **kw) -> "SpectralTransform":
        """SpectralTransform(grid; truncation, dealiasing, one_more_degree) (:305-320)."""
        from .grids import get_truncation
        truncation = truncation if truncation > 0 else get_truncation(grid.nlat_half, dealiasing)
        return cls(Spectrum(truncation + int(one_more_degree), truncation), grid, **kw)

    @classmethod
    def from_spectrum(cls, spectrum: Spectrum, Grid=OctahedralGaussianGrid, nlat_half: int = 0,
                      dealiasing: float = 2.0, **kw) -> "SpectralTransform":
        """SpectralTransform(spectrum; Grid, nlat_half, dealiasing) (:323-334)."""
        nlat_half = nlat_half if nlat_half > 0 else get_nlat_half(spectrum.mmax, dealiasing)
        return cls(spectrum, Grid(nlat_half), **kw)

    def autotune(self, K: int | None = None) -> dict:
        """Times the plan's FFT factorisation policies and launch orders on K columns and
        keeps the fastest (sht_plan_autotune); returns the choice."""
        K = self.nlayers if K is None else int(K)
        check(lib().sht_plan_autotune(self._plan, K, self._stream()), "sht_plan_autotune")
        return self.tuning()

    FOURIER = {"radix": 0, "dft32": 1, "dft16": 2}

    def set_tuning(self, mfma_min=17, eights=True, odd_first=False, legendre_xcd_groups=True, fft_pairs=True,
                   fourier="radix", dft_cols=4, legendre_light_first=False):
        """Pins the Fourier kernel (the LDS FFT with its factorisation policy: a first-pass
        MFMA prime from mfma_min on, powers of two as 4s, 8s or 16s (eights = 0, 1, 2), small
        odd radices first or last; or the ring
        DFT on MFMA with dft_cols columns per workgroup) and the launch orders: Legendre tiles
        of one order grouped on an XCD, the lightest tiles beyond the resident workgroups
        launched first, the FFT synthesis's column groups of a ring on one XCD (fft_pairs;
        sht_plan_set_tuning)."""
        leg = int(bool(legendre_xcd_groups)) | (2 if legendre_light_first else 0)
        v = (C.c_int32 * 7)(int(mfma_min), int(eights), int(bool(odd_first)), leg,
                            int(bool(fft_pairs)), self.FOURIER[fourier], int(dft_cols))
        check(lib().sht_plan_set_tuning(self._plan, v), "sht_plan_set_tuning")

    def tuning(self) -> dict:
        out = (C.c_int32 * 8)()
        check(lib().sht_plan_tuning(self._plan, out), "sht_plan_tuning")
        kinds = {v: k for k, v in self.FOURIER.items()}
        return {"tuned": bool(out[0]), "mfma_min": out[1], "eights": out[2], "odd_first": bool(out[3]),
                "legendre_xcd_groups": bool(out[4] & 1), "legendre_light_first": bool(out[4] & 2),
                "fft_pairs": bool(out[5]), "fourier": kinds[out[6]], "dft_cols": out[7]}

    def close(self):
        if getattr(self, "_plan", None) and self._plan.value:
            lib().sht_plan_destroy(self._plan)
            self._plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def norm_sphere(self) -> float:
        return 2.0 * 3.141592653589793 ** 0.5

    def __repr__(self):
        return (f"SpectralTransform({self.spectrum}, {self.grid}, nlayers={self.nlayers}, "
                f"precision={self.precision!r})")

    # -------------------------------------------------------------- checks
    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _check_tensor(self, t: torch.Tensor, n: int, dtype, what: str):
        if not t.is_cuda or t.device.index != self.device:
            raise ValueError(f"{what} must be on cuda:{self.device} (the MI355X engine has no CPU path)")
        if t.dtype != dtype:
            raise TypeError(f"{what} has dtype {t.dtype}, transform needs {dtype}")
        t2 = t if t.dim() == 2 else t.unsqueeze(0)
        if t2.stride(-1) != 1 or (t2.shape[0] > 1 and t2.stride(0) < n):
            raise ValueError(f"{what} must be contiguous along its last dimension")
        K = t2.shape[0]
        ld = t2.stride(0) if K > 1 else n
        return t2, K, ld

    def ismatching(self, x) -> bool:
        """Architectures.ismatching (spectral_transform.jl:387-403)."""
        if isinstance(x, LowerTriangularArray):
            return x.spectrum == self.spectrum and x.nlayers <= self.nlayers
        if isinstance(x, Field):
            return x.grid == self.grid and x.nlayers <= self.nlayers
        return False


def save_tuning_cache(path: str) -> None:
    """Writes every autotuner choice made or loaded in this process to `path` (one JSON
    object per line; sht_tune_cache_save)."""
    check(lib().sht_tune_cache_save(str(path).encode()), "sht_tune_cache_save")


def load_tuning_cache(path: str) -> int:
    """Merges the choices in `path` into the process's autotuner cache, so autotune() of the
    same size class applies them without timing (sht_tune_cache_load); returns how many
    loaded."""
    n = C.c_int32()
    check(lib().sht_tune_cache_load(str(path).encode(), C.byref(n)), "sht_tune_cache_load")
    return n.value


class MatrixSpectralTransform(SpectralTransform):
    """Dense matrix transform (MatrixSpectralTransform, SpeedyTransforms/src/
    matrix_transform.jl:56-260): spectral <-> grid as one FP32 GEMM each with precomputed
    matrices; the same API as SpectralTransform (the reference's GPU default up to T64)."""

    def _create(self):
        check(lib().sht_plan_create_matrix(C.byref(self._desc), C.byref(self._plan)), "sht_plan_create_matrix")


def _mismatch(S: SpectralTransform, x) -> DimensionMismatch:
    if isinstance(x, LowerTriangularArray):
        return DimensionMismatch(f"SpectralTransform for {S.spectrum.lmax}x{S.spectrum.mmax}x{S.nlayers} "
                                 f"LowerTriangularArrays with {x.spectrum}, K={x.nlayers} do not match.")
    return DimensionMismatch(f"SpectralTransform for {S.grid} x {S.nlayers} with field on {x.grid}, "
                             f"K={x.nlayers} do not match.")


def transform_(out, inp, S: SpectralTransform, unscale_coslat: bool = False):
    """In-place transform (transform!): Field <- LowerTriangularArray is spectral->grid
    (optionally dividing by cos(lat) on the fly), LowerTriangularArray <- Field is
    grid->spectral.  Returns `out`."""
    if isinstance(out, Field) and isinstance(inp, LowerTriangularArray):
        for x in (out, inp):
            if not S.ismatching(x):
                raise _mismatch(S, x)
        if out.nlayers != inp.nlayers:
            raise DimensionMismatch(f"field K={out.nlayers} and coeffs K={inp.nlayers} differ")
        coeffs = inp.data
        if coeffs.dtype == torch.complex128:  # spec->grid accepts Complex{Float64} coefficients
            coeffs = coeffs.to(torch.complex64)
        c2, K, ld_spec = S._check_tensor(coeffs, S.spectrum.nharm, torch.complex64, "coeffs")
        f2, _, ld_grid = S._check_tensor(out.data, S.grid.npoints, torch.float32, "field")
        check(lib().sht_synthesis(S._plan, c2.data_ptr(), ld_spec, f2.data_ptr(), ld_grid, K,
                                  int(bool(unscale_coslat)), S._stream()), "sht_synthesis")
        return out
    if isinstance(out, LowerTriangularArray) and isinstance(inp, Field):
        if unscale_coslat:
            raise TypeError("unscale_coslat applies to spectral->grid transforms only")
        for x in (out, inp):
            if not S.ismatching(x):
                raise _mismatch(S, x)
        if out.nlayers != inp.nlayers:
            raise DimensionMismatch(f"coeffs K={out.nlayers} and field K={inp.nlayers} differ")
        f2, K, ld_grid = S._check_tensor(inp.data, S.grid.npoints, torch.float32, "field")
        c2, _, ld_spec = S._check_tensor(out.data, S.spectrum.nharm, torch.complex64, "coeffs")
        check(lib().sht_analysis(S._plan, f2.data_ptr(), ld_grid, c2.data_ptr(), ld_spec, K,
                                 S._stream()), "sht_analysis")
        return out
    raise TypeError("transform_ needs (Field, LowerTriangularArray) or (LowerTriangularArray, Field)")


def transform(x, S: SpectralTransform, unscale_coslat: bool = False):
    """Allocating transform (spectral_transform.jl:638-682)."""
    if isinstance(x, LowerTriangularArray):
        shape = (S.grid.npoints,) if x.data.dim() == 1 else (x.nlayers, S.grid.npoints)
        out = Field(torch.empty(shape, dtype=torch.float32, device=x.data.device), S.grid)
        return transform_(out, x, S, unscale_coslat=unscale_coslat)
    if isinstance(x, Field):
        shape = (S.spectrum.nharm,) if x.data.dim() == 1 else (x.nlayers, S.spectrum.nharm)
        out = LowerTriangularArray(torch.empty(shape, dtype=torch.complex64, device=x.data.device), S.spectrum)
        return transform_(out, x, S)
    raise TypeError("transform needs a Field or a LowerTriangularArray")


def spec2grid_pullback_(coeffs_bar: LowerTriangularArray, field_bar: Field, S: SpectralTransform,
                        unscale_coslat: bool = False, accumulate: bool = True):
    """Cotangent of the coefficients of the spectral->grid transform (spec2grid_pullback!,
    SpeedyTransforms/ext/SpeedyTransformsEnzymeExt.jl:263-290): coeffs_bar += S^T field_bar
    (accumulate=False overwrites).  Returns coeffs_bar."""
    for x in (coeffs_bar, field_bar):
        if not S.ismatching(x):
            raise _mismatch(S, x)
    if coeffs_bar.nlayers != field_bar.nlayers:
        raise DimensionMismatch(f"coeffs K={coeffs_bar.nlayers} and field K={field_bar.nlayers} differ")
    f2, K, ld_grid = S._check_tensor(field_bar.data, S.grid.npoints, torch.float32, "field")
    c2, _, ld_spec = S._check_tensor(coeffs_bar.data, S.spectrum.nharm, torch.complex64, "coeffs")
    check(lib().sht_spec2grid_pullback(S._plan, f2.data_ptr(), ld_grid, c2.data_ptr(), ld_spec, K,
                                       int(bool(unscale_coslat)), int(bool(accumulate)), S._stream()),
          "sht_spec2grid_pullback")
    return coeffs_bar


def grid2spec_pullback_(field_bar: Field, coeffs_bar: LowerTriangularArray, S: SpectralTransform,
                        accumulate: bool = True):
    """Cotangent of the grid of the grid->spectral transform (grid2spec_pullback!,
    SpeedyTransforms/ext/SpeedyTransformsEnzymeExt.jl:294-319): field_bar += A^T coeffs_bar
    (accumulate=False overwrites).  Returns field_bar."""
    for x in (coeffs_bar, field_bar):
        if not S.ismatching(x):
            raise _mismatch(S, x)
    if coeffs_bar.nlayers != field_bar.nlayers:
        raise DimensionMismatch(f"coeffs K={coeffs_bar.nlayers} and field K={field_bar.nlayers} differ")
    c2, K, ld_spec = S._check_tensor(coeffs_bar.data, S.spectrum.nharm, torch.complex64, "coeffs")
    f2, _, ld_grid = S._check_tensor(field_bar.data, S.grid.npoints, torch.float32, "field")
    check(lib().sht_grid2spec_pullback(S._plan, c2.data_ptr(), ld_spec, f2.data_ptr(), ld_grid, K,
                                       int(bool(accumulate)), S._stream()), "sht_grid2spec_pullback")
    return field_bar


def _spectral_args(S: SpectralTransform, arrays, names):
    for x in arrays:
        if not S.ismatching(x):
            raise _mismatch(S, x)
    K = arrays[0].nlayers
    if any(x.nlayers != K for x in arrays):
        raise DimensionMismatch("spectral operator arrays must have the same number of layers")
    ts = [S._check_tensor(x.data, S.spectrum.nharm, torch.complex64, nm) for x, nm in zip(arrays, names)]
    if len({t[2] for t in ts}) != 1:
        raise ValueError("spectral operator arrays must share one column stride")
    return [t[0].data_ptr() for t in ts], K, ts[0][2]


def divergence_(div: LowerTriangularArray, u: LowerTriangularArray, v: LowerTriangularArray, S: SpectralTransform,
                flipsign: bool = False, add: bool = False, radius: float = 1.0):
    """div = ∇⋅(u, v) for components carrying 1/cos(lat) (divergence!,
    SpeedyTransforms/src/spectral_gradients.jl:44-125).  Returns div."""
    (pu, pv, po), K, ld = _spectral_args(S, (u, v, div), ("u", "v", "div"))
    check(lib().sht_divergence(S._plan, pu, pv, po, ld, K, 0, int(flipsign), int(add), float(radius), S._stream()),
          "sht_divergence")
    return div


def curl_(curl: LowerTriangularArray, u: LowerTriangularArray, v: LowerTriangularArray, S: SpectralTransform,
          flipsign: bool = False, add: bool = False, radius: float = 1.0):
    """curl = ∇×(u, v) for components carrying 1/cos(lat) (curl!, spectral_gradients.jl:22-41).
    Returns curl."""
    (pu, pv, po), K, ld = _spectral_args(S, (u, v, curl), ("u", "v", "curl"))
    check(lib().sht_divergence(S._plan, pu, pv, po, ld, K, 1, int(flipsign), int(add), float(radius), S._stream()),
          "sht_divergence")
    return curl


def laplace_(out: LowerTriangularArray, alms: LowerTriangularArray, S: SpectralTransform, add: bool = False,
             flipsign: bool = False, inverse: bool = False, radius: float = 1.0):
    """out = ∇² alms, or ∇⁻² alms with inverse=True (∇²! / ∇⁻²!, spectral_gradients.jl:385-466).
    May run in place.  Returns out."""
    (pa, po), K, ld = _spectral_args(S, (alms, out), ("alms", "out"))
    check(lib().sht_laplace(S._plan, pa, po, ld, K, int(inverse), int(flipsign), int(add), float(radius),
                            S._stream()), "sht_laplace")
    return out


def inverse_laplace_(out: LowerTriangularArray, alms: LowerTriangularArray, S: SpectralTransform, add: bool = False,
                     flipsign: bool = False, radius: float = 1.0):
    """∇⁻²! (spectral_gradients.jl:456-466)."""
    return laplace_(out, alms, S, add=add, flipsign=flipsign, inverse=True, radius=radius)


def transform_laplace_(out, inp, S: SpectralTransform, inverse: bool = False, flipsign: bool = False,
                       add: bool = False, radius: float = 1.0, unscale_coslat: bool = False):
    """A transform fused with the spectral Laplacian (one call, the Laplacian applied in the
    Legendre stage; the reference composes transform! and ∇²! / ∇⁻²!,
    spectral_gradients.jl:385-466):
    Field <- LowerTriangularArray: out = transform(∇² inp) (or ∇⁻² with inverse=True, negated
    with flipsign, optionally divided by cos(lat));
    LowerTriangularArray <- Field: out = ∇² transform(inp), with add and flipsign as in
    laplace_ (add accumulates into out, then the radius scaling applies).  Returns out."""
    if isinstance(out, Field) and isinstance(inp, LowerTriangularArray):
        if add:
            raise TypeError("add applies to grid->spectral transforms only")
        for x in (out, inp):
            if not S.ismatching(x):
                raise _mismatch(S, x)
        if out.nlayers != inp.nlayers:
            raise DimensionMismatch(f"field K={out.nlayers} and coeffs K={inp.nlayers} differ")
        c2, K, ld_spec = S._check_tensor(inp.data, S.spectrum.nharm, torch.complex64, "coeffs")
        f2, _, ld_grid = S._check_tensor(out.data, S.grid.npoints, torch.float32, "field")
        check(lib().sht_synthesis_laplace(S._plan, c2.data_ptr(), ld_spec, f2.data_ptr(), ld_grid, K, int(inverse),
                                          int(flipsign), float(radius), int(bool(unscale_coslat)), S._stream()),
              "sht_synthesis_laplace")
        return out
    if isinstance(out, LowerTriangularArray) and isinstance(inp, Field):
        if unscale_coslat:
            raise TypeError("unscale_coslat applies to spectral->grid transforms only")
        for x in (out, inp):
            if not S.ismatching(x):
                raise _mismatch(S, x)
        if out.nlayers != inp.nlayers:
            raise DimensionMismatch(f"coeffs K={out.nlayers} and field K={inp.nlayers} differ")
        f2, K, ld_grid = S._check_tensor(inp.data, S.grid.npoints, torch.float32, "field")
        c2, _, ld_spec = S._check_tensor(out.data, S.spectrum.nharm, torch.complex64, "coeffs")
        check(lib().sht_analysis_laplace(S._plan, f2.data_ptr(), ld_grid, c2.data_ptr(), ld_spec, K, int(inverse),
                                         int(flipsign), int(add), float(radius), S._stream()),
              "sht_analysis_laplace")
        return out
    raise TypeError(f"transform_laplace_({type(out).__name__}, {type(inp).__name__}) is not defined")


def gradient_(dpdx: LowerTriangularArray, dpdy: LowerTriangularArray, p: LowerTriangularArray, S: SpectralTransform,
              radius: float = 1.0):
    """dpdx = ∂p/∂λ, dpdy = cos(lat) ∂p/∂φ (∇!, spectral_gradients.jl:483-527).  Returns (dpdx, dpdy)."""
    (pp, px, py), K, ld = _spectral_args(S, (p, dpdx, dpdy), ("p", "dpdx", "dpdy"))
    check(lib().sht_gradient(S._plan, pp, px, py, ld, K, float(radius), S._stream()), "sht_gradient")
    return dpdx, dpdy


def UV_from_vordiv_(U: LowerTriangularArray, V: LowerTriangularArray, vor: LowerTriangularArray,
                    div: LowerTriangularArray, S: SpectralTransform, radius: float = 1.0):
    """U, V = (u, v) cos(lat) from vorticity and divergence (spectral_gradients.jl:309-366)."""
    for x in (U, V, vor, div):
        if not S.ismatching(x):
            raise _mismatch(S, x)
    K = vor.nlayers
    if not (U.nlayers == V.nlayers == div.nlayers == K):
        raise DimensionMismatch("U, V, vor, div must have the same number of layers")
    ts = [S._check_tensor(x.data, S.spectrum.nharm, torch.complex64, nm)
          for x, nm in ((vor, "vor"), (div, "div"), (U, "U"), (V, "V"))]
    lds = {t[2] for t in ts}
    if len(lds) != 1:
        raise ValueError("U, V, vor, div must share one column stride")
    check(lib().sht_uv_from_vordiv(S._plan, ts[0][0].data_ptr(), ts[1][0].data_ptr(), ts[2][0].data_ptr(),
                                   ts[3][0].data_ptr(), ts[0][2], K, float(radius), S._stream()),
          "sht_uv_from_vordiv")
    return U, V


class SpectralBatch:
    """An uneven batch of independent transforms (any mix of SpectralTransforms, layer counts
    and directions) run as one kernel launch per stage, replayed from a hipGraph.

    items: sequences (S, inp, out) for a single direction -- (S, LowerTriangularArray, Field)
    is spectral->grid, (S, Field, LowerTriangularArray) grid->spectral -- or (S, coeffs_in,
    coeffs_out, mid_field) for a round trip; an optional trailing bool is unscale_coslat.
    The tensors are bound at construction: every launch() repeats the same transforms."""

    def __init__(self, items, graph: bool = True, force_graph_failure: bool = False):
        arr = (_lib.ShtBatchItem * len(items))()
        self._keep = []
        self.items = list(items)
        for i, it in enumerate(items):
            S, inp, out = it[0], it[1], it[2]
            rest = list(it[3:])
            mid = rest.pop(0) if rest and isinstance(rest[0], Field) else None
            unscale = bool(rest[0]) if rest else False
            Ks = []
            if mid is not None:
                direction = _lib.SHT_DIR_ROUNDTRIP
                a, K, ld_in = S._check_tensor(inp.data, S.spectrum.nharm, torch.complex64, "coeffs")
                m2, Km, ld_mid = S._check_tensor(mid.data, S.grid.npoints, torch.float32, "field")
                o, Ko, ld_out = S._check_tensor(out.data, S.spectrum.nharm, torch.complex64, "coeffs")
                checks, Ks = (inp, mid, out), [Km, Ko]
            elif isinstance(inp, LowerTriangularArray):
                direction = _lib.SHT_DIR_SYNTHESIS
                a, K, ld_in = S._check_tensor(inp.data, S.spectrum.nharm, torch.complex64, "coeffs")
                o, Ko, ld_out = S._check_tensor(out.data, S.grid.npoints, torch.float32, "field")
                m2, ld_mid, checks, Ks = None, 0, (inp, out), [Ko]
            else:
                direction = _lib.SHT_DIR_ANALYSIS
                a, K, ld_in = S._check_tensor(inp.data, S.grid.npoints, torch.float32, "field")
                o, Ko, ld_out = S._check_tensor(out.data, S.spectrum.nharm, torch.complex64, "coeffs")
                m2, ld_mid, checks, Ks = None, 0, (inp, out), [Ko]
            for x in checks:
                if not S.ismatching(x):
                    raise _mismatch(S, x)
            if any(k != K for k in Ks):  # the kernels write K columns of every output
                raise DimensionMismatch(f"batch item {i}: input K={K}, output/grid K={Ks} differ")
            arr[i] = _lib.ShtBatchItem(S._plan.value, direction, K, int(unscale), 0, a.data_ptr(), ld_in,
                                       m2.data_ptr() if m2 is not None else None, ld_mid, o.data_ptr(), ld_out)
            self._keep.append((S, a, m2, o))
        flags = (0 if graph else _lib.SHT_BATCH_NO_GRAPH) | (
            _lib.SHT_BATCH_FORCE_GRAPH_FAILURE if force_graph_failure else 0)
        self._batch = C.c_void_p()
        self.device = items[0][0].device
        check(lib().sht_batch_create(arr, len(items), flags, C.byref(self._batch)), "sht_batch_create")

    @property
    def uses_graph(self) -> bool:
        return bool(lib().sht_batch_uses_graph(self._batch))

    def launch(self):
        check(lib().sht_batch_launch(self._batch, C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)),
              "sht_batch_launch")

    def close(self):
        if getattr(self, "_batch", None) and self._batch.value:
            lib().sht_batch_destroy(self._batch)
            self._batch = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
