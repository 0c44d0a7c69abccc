# Julia binding of libsht for SpeedyTransforms: a package extension that routes
# transform! on AMD GPUs through the MI355X engine.
#
# It adds methods for the two positional cores every transform! call goes through,
# SpeedyTransforms._transform_grid! and ._transform_spec!
# (SpeedyTransforms/src/spectral_transform.jl:533,597), for Float32 transforms on
# GPU{ROCBackend}.  The SpectralTransform type is unchanged; the engine's plan for each
# transform lives in a global IdDict keyed by the transform (the same pattern the
# reference uses for its per-transform GPU caches, ext/gpu_graphs_common.jl:116-121).
# SpectralTransform is immutable, so no finalizer can be attached to it: plans live until
# free_plans() (also registered with atexit) destroys them.  libsht keeps a Fourier scratch
# per (plan, stream), so tasks on different AMDGPU streams may transform with one S at once;
# the reference's ScratchMemory argument is therefore not needed and is ignored.
#
# Wiring into SpeedyTransforms/Project.toml:
#   [weakdeps]    AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
#   [extensions]  SpeedyTransformsMI355XExt = "AMDGPU"
# and LIBSHT set (environment variable) to the absolute path of libsht.so.
#
# Precision: ENV["LIBSHT_PRECISION"] in ("fp32", "fp16x3", "bf16x3"), default "fp32".
module SpeedyTransformsMI355XExt

import AMDGPU
using SpeedyTransforms
using SpeedyTransforms.RingGrids
using SpeedyTransforms.LowerTriangularArrays
import SpeedyTransforms: SpectralTransform, _transform_grid!, _transform_spec!, ScratchMemory
import SpeedyWeatherInternals.Architectures: GPU, ismatching

const libsht = get(ENV, "LIBSHT", "libsht.so")

# mirrors struct sht_desc of include/sht.h
struct ShtDesc
    grid::Int32
    nlat_half::Int32
    lmax::Int32
    mmax::Int32
    nlayers::Int32
    shortcut::Int32
    precision::Int32
    device::Int32
end

const SHT_OK = Int32(0)
const SHT_ERR_DIM = Int32(2)

const PLANS = IdDict{Any, Ptr{Cvoid}}()
const PLANS_LOCK = ReentrantLock()

grid_code(::OctahedralGaussianGrid) = Int32(0)
grid_code(::FullGaussianGrid) = Int32(1)
grid_code(g) = throw(ArgumentError("libsht supports OctahedralGaussianGrid and FullGaussianGrid, not $(typeof(g))"))

function shortcut_code(S)
    s = S.LegendreShortcut
    s === SpeedyTransforms.LegendreShortcutLinear && return Int32(0)
    s === SpeedyTransforms.LegendreShortcutQuadratic && return Int32(1)
    s === SpeedyTransforms.LegendreShortcutCubic && return Int32(2)
    s === SpeedyTransforms.LegendreShortcutLinQuadCoslat² && return Int32(3)
    s === SpeedyTransforms.LegendreShortcutLinCubCoslat && return Int32(4)
    throw(ArgumentError("libsht implements the five Legendre shortcuts of legendre_shortcuts.jl"))
end

function precision_code()
    p = get(ENV, "LIBSHT_PRECISION", "fp32")
    p == "fp32" && return Int32(0)
    p == "fp16x3" && return Int32(1)
    p == "bf16x3" && return Int32(2)
    throw(ArgumentError("LIBSHT_PRECISION must be fp32, fp16x3 or bf16x3"))
end

last_error() = unsafe_string(ccall((:sht_last_error, libsht), Cstring, ()))

function check(rc::Int32, S, x)
    rc == SHT_OK && return nothing
    rc == SHT_ERR_DIM && throw(DimensionMismatch(S, x))
    error("libsht: " * last_error())
end

function plan(S::SpectralTransform)
    lock(PLANS_LOCK) do
        get!(PLANS, S) do
            desc = Ref(ShtDesc(grid_code(S.grid), S.grid.nlat_half, S.spectrum.lmax, S.spectrum.mmax,
                               S.nlayers, shortcut_code(S), precision_code(), AMDGPU.device_id(AMDGPU.device()) - 1))
            p = Ref{Ptr{Cvoid}}(C_NULL)
            rc = ccall((:sht_plan_create, libsht), Cint, (Ref{ShtDesc}, Ref{Ptr{Cvoid}}), desc, p)
            rc == SHT_OK || error("libsht plan: " * last_error())
            p[]
        end
    end
end

"""Destroys every plan made so far (waits for the device first); later transforms make new ones."""
function free_plans()
    lock(PLANS_LOCK) do
        for p in values(PLANS)
            ccall((:sht_plan_destroy, libsht), Cint, (Ptr{Cvoid},), p)
        end
        empty!(PLANS)
    end
    return nothing
end

__init__() = atexit(free_plans)

stream_handle() = AMDGPU.stream().stream
ncols(A) = ndims(A) == 1 ? 1 : size(A, 2)
colstride(A, n) = ndims(A) == 1 ? n : stride(A, 2)

const GPUSpectralTransform32 = SpectralTransform{Float32, <:GPU{<:AMDGPU.ROCBackend}}

# spectral -> grid (spectral_transform.jl:533-570)
function _transform_grid!(field::AbstractField, coeffs::LowerTriangularArray,
                          scratch_memory::ScratchMemory, S::GPUSpectralTransform32, unscale_coslat::Bool)
    ismatching(S, field) || throw(DimensionMismatch(S, field))
    ismatching(S, coeffs) || throw(DimensionMismatch(S, coeffs))
    c = coeffs.data
    eltype(c) == ComplexF32 || (c = ComplexF32.(c))    # Complex{Float64} input is accepted
    f = field.data
    eltype(f) == Float32 || throw(ArgumentError("field must be Float32 like the transform"))
    K = ncols(f)
    ncols(c) == K || throw(DimensionMismatch(S, coeffs))   # the kernels read K columns of coeffs
    GC.@preserve c f begin
        rc = ccall((:sht_synthesis, libsht), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int32, Int32, Ptr{Cvoid}),
                   plan(S), pointer(c), colstride(c, size(c, 1)), pointer(f), colstride(f, size(f, 1)),
                   K, Int32(unscale_coslat), stream_handle())
        check(rc, S, field)
    end
    return field
end

# grid -> spectral (spectral_transform.jl:597-631)
function _transform_spec!(coeffs::LowerTriangularArray, field::AbstractField,
                          scratch_memory::ScratchMemory, S::GPUSpectralTransform32)
    ismatching(S, field) || throw(DimensionMismatch(S, field))
    ismatching(S, coeffs) || throw(DimensionMismatch(S, coeffs))
    c = coeffs.data
    f = field.data
    K = ncols(f)
    ncols(c) == K || throw(DimensionMismatch(S, coeffs))   # the kernels write K columns of coeffs
    GC.@preserve c f begin
        rc = ccall((:sht_analysis, libsht), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int64, Int32, Ptr{Cvoid}),
                   plan(S), pointer(f), colstride(f, size(f, 1)), pointer(c), colstride(c, size(c, 1)),
                   K, stream_handle())
        check(rc, S, coeffs)
    end
    return coeffs
end

end # module
