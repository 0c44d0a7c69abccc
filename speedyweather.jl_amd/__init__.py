"""MI355X spherical harmonic transform engine for SpeedyWeather (host-side mirror of
SpeedyTransforms' API over libsht's C ABI; hand-written gfx950 kernels underneath)."""
from ._lib import DimensionMismatch, SHTError, lib
from .grids import (Field, FullGaussianGrid, OctahedralGaussianGrid, get_nlat_half, get_truncation,
                    roundup_fft)
from . import dycore, learned
from .mlp import ColumnMLP
from .spectrum import LowerTriangularArray, Spectrum
from .synthetic import seeded_grid, seeded_spectrum
from .transform import (MatrixSpectralTransform, SpectralBatch, SpectralTransform, UV_from_vordiv_, curl_, divergence_,
                        gradient_, grid2spec_pullback_, inverse_laplace_, laplace_, load_tuning_cache,
                        save_tuning_cache, spec2grid_pullback_, transform, transform_, transform_laplace_)

__all__ = [
    "ColumnMLP", "MatrixSpectralTransform", "DimensionMismatch", "SHTError", "lib", "Field", "FullGaussianGrid", "OctahedralGaussianGrid",
    "get_nlat_half", "get_truncation", "roundup_fft", "LowerTriangularArray", "Spectrum",
    "seeded_grid", "seeded_spectrum", "SpectralBatch", "SpectralTransform", "UV_from_vordiv_", "transform",
    "spec2grid_pullback_", "grid2spec_pullback_", "divergence_", "curl_", "laplace_", "inverse_laplace_",
    "gradient_",
    "transform_", "transform_laplace_", "save_tuning_cache", "load_tuning_cache",
]
