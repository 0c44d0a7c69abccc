/* libsht: spherical harmonic transform engine for SpeedyWeather on AMD MI355X (gfx950).
 *
 * C ABI (plain pointers and sizes, no C++ or torch types) for the hot path of
 * SpeedyTransforms' transform!:
 *
 *   spectral -> grid  transform!(field, coeffs, S; unscale_coslat)
 *                     SpeedyTransforms/src/spectral_transform.jl:515-570
 *   grid -> spectral  transform!(coeffs, field, S)
 *                     SpeedyTransforms/src/spectral_transform.jl:579-631
 *
 * A host binding (Julia ccall, Python ctypes) overrides the positional cores
 * _transform_grid! / _transform_spec! (spectral_transform.jl:533,597) with
 * sht_synthesis / sht_analysis; a plan replaces the device-side state of one
 * SpectralTransform (spectral_transform.jl:118-302).
 *
 * Conventions (same as the reference):
 *   - spectral data: LowerTriangularArray layout (LowerTriangularArrays/src/
 *     lower_triangular_array.jl:367): column m (0-based) of lmax-m degrees, m-major,
 *     complex float interleaved; column k of K starts at coeffs + k*ld_spec (complex
 *     elements).
 *   - grid data: RingGrids Field layout (RingGrids/src/field.jl:7-18): rings north to
 *     south, each from 0 degrees east; column k starts at field + k*ld_grid (floats).
 *   - all data pointers passed to sht_synthesis / sht_analysis are DEVICE pointers.
 *   - outputs are overwritten, inputs are not modified.
 *
 * Every function returns an int status (SHT_OK = 0); nothing throws across the ABI.
 * sht_last_error() returns a message for the last failure on the calling thread.
 */
#ifndef SHT_H
#define SHT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SHT_OK 0
#define SHT_ERR_ARG 1          /* invalid argument (null pointer, bad enum) */
#define SHT_ERR_DIM 2          /* size mismatch: DimensionMismatch(S, coeffs|field) */
#define SHT_ERR_TYPE 3         /* number format mismatch */
#define SHT_ERR_HIP 4          /* HIP runtime error */
#define SHT_ERR_UNSUPPORTED 5  /* configuration the engine does not implement */
#define SHT_ERR_ALLOC 6        /* out of device or host memory */

/* grid kinds (RingGrids/src/grids/octahedral_gaussian.jl, full_gaussian.jl) */
#define SHT_GRID_OCTAHEDRAL_GAUSSIAN 0
#define SHT_GRID_FULL_GAUSSIAN 1

/* Legendre shortcuts (SpeedyTransforms/src/legendre_shortcuts.jl:21-49): the last order
 * kept on a ring, mmax_truncation(ring) = min(s(nlon, latd), mmax - 1) with s = nlon ÷ 2,
 * nlon ÷ 3, nlon ÷ 4, floor(nlon / (2 + cosd(latd)^2)), floor(nlon / (2 + 2 cosd(latd))).
 * The last two need not be monotonic in the ring index (a ring nearer the equator may keep
 * fewer orders than one nearer the pole); every ring still contributes exactly its own
 * orders 0..mmax_truncation(ring). */
#define SHT_SHORTCUT_LINEAR 0
#define SHT_SHORTCUT_QUADRATIC 1
#define SHT_SHORTCUT_CUBIC 2
#define SHT_SHORTCUT_LINQUAD_COSLAT2 3
#define SHT_SHORTCUT_LINCUB_COSLAT 4

/* arithmetic of the Legendre contraction (FFT and all other steps are FP32) */
#define SHT_PREC_FP32 0      /* FP32 products, FP32 accumulation (MFMA f32) */
#define SHT_PREC_FP16X3 1    /* scaled FP16 hi/lo split, 3 products, FP32 accumulation */
#define SHT_PREC_BF16X3 2    /* three-term BF16 split (hi+mid+lo), 6 products, FP32 accumulation */

typedef struct sht_desc {
    int32_t grid;        /* SHT_GRID_* */
    int32_t nlat_half;   /* rings per hemisphere */
    int32_t lmax;        /* 1-based number of degrees (Spectrum.lmax) */
    int32_t mmax;        /* 1-based number of orders (Spectrum.mmax) */
    int32_t nlayers;     /* capacity: largest K any call may pass (SpectralTransform.nlayers) */
    int32_t shortcut;    /* SHT_SHORTCUT_* */
    int32_t precision;   /* SHT_PREC_* */
    int32_t device;      /* HIP device ordinal the plan lives on */
} sht_desc;

typedef struct sht_plan sht_plan;

/* Sizes implied by a descriptor (host only, no GPU needed). */
typedef struct sht_sizes {
    int64_t npoints;     /* grid points per column */
    int64_t nharm;       /* spectral coefficients per column */
    int32_t nlat;        /* rings */
    int32_t nlon_max;    /* points on the longest ring */
    int64_t nfreq_total; /* sum over northern rings of nlon/2+1 (Fourier buffer rows) */
    int64_t device_bytes;/* device memory the plan allocates */
} sht_sizes;

/* Precompute a plan (SpectralTransform(spectrum, grid; NF=Float32, nlayers,
 * LegendreShortcut), spectral_transform.jl:118-302): Gaussian latitudes and weights and the
 * Legendre polynomials in FP64, stored in the layout the precision mode needs; FFT
 * factorisations and twiddles; scratch for nlayers columns.  Blocking, host side.
 * Scratch is per (plan, stream): the null stream's is allocated here, another stream's on
 * its first call (at most 16 streams per plan), so calls on different streams may overlap
 * on one plan, as the reference's explicit-scratch transform! permits
 * (spectral_transform.jl:522-531). */
int sht_plan_create(const sht_desc *desc, sht_plan **plan);
int sht_plan_destroy(sht_plan *plan);
int sht_plan_sizes(const sht_desc *desc, sht_sizes *sizes);

/* spectral -> grid, replaces _transform_grid!(field, coeffs, scratch, S, unscale_coslat)
 * (spectral_transform.jl:533-570).  coeffs: K columns of nharm complex floats (stride
 * ld_spec >= nharm); field: K columns of npoints floats (stride ld_grid >= npoints).
 * Asynchronous on `stream` (a hipStream_t; 0 = null stream). */
int sht_synthesis(sht_plan *plan, const void *coeffs, int64_t ld_spec, void *field,
                  int64_t ld_grid, int32_t K, int32_t unscale_coslat, void *stream);

/* Dense matrix transform, replaces MatrixSpectralTransform(spectrum, grid; NF, nlayers)
 * (SpeedyTransforms/src/matrix_transform.jl:56-260; the GPU default for truncations up to 64,
 * SpeedyWeather/src/dynamics/spectral_grid.jl:343-349).  Builds a regular plan plus the two
 * dense transform matrices (FP64-computed, stored FP32); sht_synthesis and sht_analysis then
 * run as one FP32 GEMM each (a split-K GEMM of this library, its partial sums in the
 * stream's scratch), every other entry point as on a regular plan.
 * SHT_ERR_UNSUPPORTED if the matrices would exceed 4 GiB. */
int sht_plan_create_matrix(const sht_desc *desc, sht_plan **plan);

/* grid -> spectral, replaces _transform_spec!(coeffs, field, scratch, S)
 * (spectral_transform.jl:597-631). */
int sht_analysis(sht_plan *plan, const void *field, int64_t ld_grid, void *coeffs,
                 int64_t ld_spec, int32_t K, void *stream);

/* Allocation-free transform (no reference counterpart): for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  direction SHT_DIR_SYNTHESIS runs sht_synthesis, SHT_DIR_ANALYSIS runs
 * sht_analysis (unscale_coslat must then be 0), same layouts, strides, streams and results;
 * the call's Fourier scratch is a static device array of SHT_STATIC_SCRATCH_BYTES that each
 * GPU holds from the library's load until the process ends (its address is looked up when
 * the first plan on the device is created), so the call allocates nothing.  The array is
 * shared by every sht_transform_static call on a GPU, whatever the plan or stream: calls
 * that overlap on one GPU race, so serialise them (one stream, or events).  Nothing in it
 * needs to start at zero (each stage writes every row the next one reads).  SHT_ERR_DIM if
 * K columns need more than the array: plan rows sum(nlon/2+1) x K x 16 bytes (T255: K <= 108),
 * or a dense plan's GEMM workspace. */
#define SHT_STATIC_SCRATCH_BYTES (64ll << 20)
int sht_transform_static(sht_plan *plan, int32_t direction, const void *in, int64_t ld_in, void *out,
                         int64_t ld_out, int32_t K, int32_t unscale_coslat, void *stream);

/* Pullbacks (adjoints) of the two transforms, replace spec2grid_pullback! and
 * grid2spec_pullback! (SpeedyTransforms/ext/SpeedyTransformsEnzymeExt.jl:263-319).
 * sht_spec2grid_pullback: coeffs_bar (+)= S^T field_bar for the synthesis S (with the same
 * unscale_coslat as the forward call): forward FFT of field_bar, times 2 on the interior
 * bins (and 1/cos lat), divided by the solid angle, forward Legendre.
 * sht_grid2spec_pullback: field_bar (+)= A^T coeffs_bar for the analysis A: inverse
 * Legendre, times the solid angle, halved on the interior bins, inverse FFT.
 * accumulate = 1 adds into the output as the reference does; 0 overwrites it.  Layouts,
 * strides and streams as in sht_synthesis / sht_analysis. */
int sht_spec2grid_pullback(sht_plan *plan, const void *field_bar, int64_t ld_grid, void *coeffs_bar,
                           int64_t ld_spec, int32_t K, int32_t unscale_coslat, int32_t accumulate, void *stream);
int sht_grid2spec_pullback(sht_plan *plan, const void *coeffs_bar, int64_t ld_spec, void *field_bar,
                           int64_t ld_grid, int32_t K, int32_t accumulate, void *stream);

/* U, V = (u, v) cos(lat) spectra from vorticity and divergence, replaces
 * UV_from_vordiv!(U, V, vor, div, S; radius) (SpeedyTransforms/src/spectral_gradients.jl:
 * 309-366).  All four arrays: K columns of nharm complex floats with stride ld. */
int sht_uv_from_vordiv(sht_plan *plan, const void *vor, const void *div, void *U, void *V,
                       int64_t ld, int32_t K, double radius, void *stream);

/* Spectral gradient operators on the plan's spectrum (unit sphere unless radius != 1), all
 * arrays K columns of nharm complex floats with stride ld; outputs must not alias inputs
 * (the Laplacian may run in place).  flipsign negates, add accumulates, the radius scaling
 * applies after adding, as in the reference.
 * sht_divergence: out = div(u, v) (curl = 0) or curl(u, v) (curl = 1) of a vector whose
 *   components carry the 1/cos(lat) scaling; replaces divergence! / curl!
 *   (SpeedyTransforms/src/spectral_gradients.jl:30-125).
 * sht_laplace: out = ∇² alms or ∇⁻² alms (inverse = 1, the l = 0 mode set to 0); replaces
 *   ∇²! / ∇⁻²! (spectral_gradients.jl:385-466).
 * sht_gradient: dpdx = ∂p/∂λ, dpdy = cos(lat) ∂p/∂φ; replaces ∇! (spectral_gradients.jl:483-527). */
int sht_divergence(sht_plan *plan, const void *u, const void *v, void *out, int64_t ld, int32_t K, int32_t curl,
                   int32_t flipsign, int32_t add, double radius, void *stream);
int sht_laplace(sht_plan *plan, const void *alms, void *out, int64_t ld, int32_t K, int32_t inverse,
                int32_t flipsign, int32_t add, double radius, void *stream);
int sht_gradient(sht_plan *plan, const void *p, void *dpdx, void *dpdy, int64_t ld, int32_t K, double radius,
                 void *stream);

/* Transforms fused with the spectral Laplacian: one call replaces a transform plus a
 * ∇²! / ∇⁻²! pass (spectral_gradients.jl:385-466) over the spectra, with the Laplacian
 * applied inside the Legendre stage.  No single reference call: they compose
 * transform!(field, ∇²(alms), S) (e.g. the streamfunction from vorticity through ∇⁻²) and
 * ∇²!(out, transform!(alms, field, S); add, flipsign) (the Bernoulli potential's ∇² into
 * the divergence tendency, SpeedyWeather/src/dynamics/tendencies.jl:1123-1212).  inverse,
 * flipsign, add and radius as in sht_laplace (add accumulates into coeffs, then the radius
 * scaling applies); layouts, strides and streams as in sht_synthesis / sht_analysis.  On a
 * dense (matrix) plan they run the regular transform kernels. */
int sht_synthesis_laplace(sht_plan *plan, const void *coeffs, int64_t ld_spec, void *field, int64_t ld_grid,
                          int32_t K, int32_t inverse, int32_t flipsign, double radius, int32_t unscale_coslat,
                          void *stream);
int sht_analysis_laplace(sht_plan *plan, const void *field, int64_t ld_grid, void *coeffs, int64_t ld_spec, int32_t K,
                         int32_t inverse, int32_t flipsign, int32_t add, double radius, void *stream);

/* Uneven batches: many independent transforms (any mix of plans, i.e. truncations and
 * grids, layer counts and fields) run as ONE launch per stage, their work tiles merged
 * and ordered heaviest first, replayed from an explicitly built hipGraph.  No reference
 * counterpart (the reference transforms one SpectralTransform per call and, on CUDA only,
 * captures the Fourier stage in a graph: SpeedyTransforms/ext/gpu_graphs_common.jl:293-331).
 * Items keep the sht_synthesis / sht_analysis conventions; a round trip writes its grid to
 * `mid` and its coefficients to `out`.  All plans must share one precision mode and device;
 * K <= the plan's nlayers.  The batch owns its Fourier scratch, so items on one plan do not
 * race.  The pointers are fixed at creation: a launch replays the same transforms. */
#define SHT_DIR_SYNTHESIS 0
#define SHT_DIR_ANALYSIS 1
#define SHT_DIR_ROUNDTRIP 2
typedef struct sht_batch_item {
    sht_plan *plan;
    int32_t direction;      /* SHT_DIR_* */
    int32_t K;
    int32_t unscale_coslat; /* synthesis and round trip */
    int32_t reserved;
    const void *in;         /* synthesis, round trip: coefficients; analysis: grid */
    int64_t ld_in;
    void *mid;              /* round trip: grid in between; otherwise unused */
    int64_t ld_mid;
    void *out;              /* synthesis: grid; analysis, round trip: coefficients */
    int64_t ld_out;
} sht_batch_item;
typedef struct sht_batch sht_batch;
#define SHT_BATCH_NO_GRAPH 1             /* flag: launch the stages directly */
#define SHT_BATCH_FORCE_GRAPH_FAILURE 2  /* flag for tests: treat graph instantiation as failed */
int sht_batch_create(const sht_batch_item *items, int32_t n, int32_t flags, sht_batch **batch);
int sht_batch_launch(sht_batch *batch, void *stream);
/* 1 if launches replay a hipGraph, 0 if they fall back to direct kernel launches */
int sht_batch_uses_graph(const sht_batch *batch);
int sht_batch_destroy(sht_batch *batch);

/* Learned column correction (an addition of this library; no reference counterpart): an MLP
 * 32 -> 128 -> 128 -> 32 with GELU per grid column (8 levels x {u, v, T, ln ps}), BF16
 * weights with FP32 master copy and FP32 accumulation, trained with Adam on the mean
 * squared error.  x, t, y, grad and loss are device pointers; x, t, y are B x 32 row-major.
 * Parameters (sht_mlp_nparams() floats): W1[128][32] b1[128] W2[128][128] b2[128]
 * W3[32][128] b3[32]. */
typedef struct sht_mlp sht_mlp;
int sht_mlp_nparams(void);
int sht_mlp_create(int32_t device, int32_t max_batch, sht_mlp **mlp);
int sht_mlp_set_params(sht_mlp *mlp, const float *host_params);
int sht_mlp_get_params(sht_mlp *mlp, float *host_params);
int sht_mlp_forward(sht_mlp *mlp, const float *x, float *y, int32_t B, void *stream);
/* gradient of the loss w.r.t. the parameters (no update) and the loss */
int sht_mlp_gradient(sht_mlp *mlp, const float *x, const float *t, int32_t B, float *grad, float *loss,
                     void *stream);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr; loss may be null */
int sht_mlp_train_step(sht_mlp *mlp, const float *x, const float *t, int32_t B, float lr, float *loss,
                       void *stream);
int sht_mlp_destroy(sht_mlp *mlp);

/* Autotuner (no reference counterpart; kernel, tile and order choices per size class): on K
 * columns of scratch data, times the Fourier stage's kernels -- the LDS ring FFT under each
 * factorisation policy (first-pass MFMA threshold 11 or 17, powers of two as 4s, 8s or 16s,
 * small odd radices first or last) and the ring DFT on MFMA (FP32, or FP16 hi/lo on split plans)
 * with 4 or 8 columns per workgroup -- then the launch orders (Legendre tiles grouped per
 * order m on one XCD or plain heaviest-first, each with or without the lightest tiles first;
 * LDS-FFT synthesis with all column groups of a ring on one XCD, or in rank order),
 * keeps the fastest in the plan and caches it per (grid, nlat_half, lmax, mmax, nlayers,
 * precision, K) for later plans in the process.  Synchronises the stream; results are
 * unchanged up to the rounding of the Fourier stage's arithmetic. */
#define SHT_FOURIER_RADIX 0   /* mixed-radix LDS FFT, FP32 */
#define SHT_FOURIER_DFT32 1   /* radix-4 pass + p-point DFT on MFMA, FP32 products */
#define SHT_FOURIER_DFT16 2   /* the same with scaled FP16 hi/lo products (split plans) */
int sht_plan_autotune(sht_plan *plan, int32_t K, void *stream);
/* out[8] = {tuned, MFMA threshold, eights, odd_first, Legendre launch order, FFT grouping,
 * Fourier kernel (SHT_FOURIER_*), DFT columns per workgroup}; the Legendre launch order is
 * bit 0: the tiles of one order m grouped on one XCD, bit 1: the lightest tiles beyond the
 * device's resident workgroups launched first (no tail of late light tiles). */
int sht_plan_tuning(const sht_plan *plan, int32_t *out);
/* Tuning cache file (one flat JSON object per line): save writes every choice the autotuner
 * made or loaded in this process; load merges a file's choices into the cache, so a later
 * sht_plan_autotune of the same size class applies them without timing (nloaded may be
 * null).  SHT_ERR_ARG on unreadable files or malformed lines (the valid lines still load). */
int sht_tune_cache_save(const char *path);
int sht_tune_cache_load(const char *path, int32_t *nloaded);
/* pins a choice: in[7] = {MFMA threshold (11 or 17), powers of two as 4s / 8s / 16s (eights =
 * 0 / 1 / 2), odd_first, Legendre launch order (0..3, bits as above), FFT grouping (1: the
 * synthesis's column groups of a ring on one XCD), Fourier kernel (SHT_FOURIER_*), DFT columns
 * per workgroup (4 or 8)}.  Every choice's tables are built at plan creation, so this
 * changes host state only: calls already queued keep what they were launched with.  Not
 * to be called concurrently with other calls on the plan from another host thread. */
int sht_plan_set_tuning(sht_plan *plan, const int32_t *in);

/* Graph replay of single calls (the reference caches a captured graph of its Fourier stage
 * per transform and buffer on CUDA: SpeedyTransforms/ext/gpu_graphs_common.jl:23,121,293-331,
 * MAX_GRAPHS = 64, capture failure -> direct loop).  With graphs on, sht_synthesis /
 * sht_analysis and the calls built on them replay an explicitly built graph of their two
 * stage kernels, one per distinct call (direction, buffers, strides, K, flags, scratch,
 * tuning), built on its first use while fewer than 64 per direction are cached; a call whose
 * graph fails to build or instantiate, calls beyond the cap, calls on a capturing stream and
 * calls while profiling launch directly, and so do a dense matrix plan's GEMM calls.  Results
 * are identical either way.  Off by default.
 * SHT_GRAPHS_FORCE_FAILURE (tests): treat every instantiation as failed. */
#define SHT_GRAPHS_FORCE_FAILURE 1
int sht_plan_set_graphs(sht_plan *plan, int32_t enable, int32_t flags);
/* graphs cached on the plan (both directions) and graphs that failed */
int sht_plan_graph_count(sht_plan *plan, int32_t *cached, int32_t *failed);

/* Per-kernel timing (no reference counterpart; the reference profiled ad hoc with
 * CUDA.@profile / BenchmarkTools, docs/dev/2026-08/legendre-gpu-optimization.md).
 * Between begin and end, the first max_calls launches of every kernel are bracketed by
 * HIP events on the call's stream; end waits for them and returns the mean duration (ms)
 * and launch count per kernel, indexed by SHT_KERNEL_*. */
#define SHT_KERNEL_LEGENDRE_SYNTHESIS 0
#define SHT_KERNEL_FFT_SYNTHESIS 1
#define SHT_KERNEL_FFT_ANALYSIS 2
#define SHT_KERNEL_LEGENDRE_ANALYSIS 3
#define SHT_NKERNELS 4
int sht_profile_begin(sht_plan *plan, int32_t max_calls);
/* Mean duration (ms) of one launch of a single stage kernel (SHT_KERNEL_*), timed with HIP
 * events around `reps` back-to-back launches on `stream` (after one untimed launch), on
 * K columns with the plan's chosen launch orders.  The Legendre synthesis reads coeffs and
 * the FFT synthesis writes field, the FFT analysis reads field and the Legendre analysis
 * writes coeffs; the Fourier stages hand over through the stream's scratch, so run a
 * transform on the stream first for meaningful values.  Blocks until the launches end. */
int sht_time_stage(sht_plan *plan, int32_t stage, void *coeffs, int64_t ld_spec, void *field, int64_t ld_grid,
                   int32_t K, int32_t reps, void *stream, double *avg_ms);
int sht_profile_end(sht_plan *plan, double *avg_ms, int32_t *counts);

/* Host-only precompute exposed for verification (no GPU needed). */
/* Gauss-Legendre nodes (ascending) and weights of degree nlat
 * (RingGrids/src/grids/full_gaussian.jl:48-50, quadrature_weights.jl:11). */
int sht_gauss_legendre(int32_t nlat, double *nodes, double *weights);
/* λ_l^m(x) for all 0 <= m < mmax, m <= l < lmax in LowerTriangularArray order
 * (the Λ column that spectral_transform.jl:167-172 stores per ring). */
int sht_legendre_column(int32_t lmax, int32_t mmax, double x, double *out);

const char *sht_last_error(void);
const char *sht_version(void);

#ifdef __cplusplus
}
#endif
#endif
