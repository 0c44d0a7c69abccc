// Internal definitions shared by the host plan code and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/sht.h"

namespace sht {

// roctx range over one host-side entry point (visible with rocprofv3 --marker-trace;
// negligible without a tool attached)
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePush(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

// Column group: 32 complex columns = 64 real MFMA columns per Legendre workgroup.
constexpr int kColsComplex = 32;
constexpr int kColsReal = 64;
// Rings per synthesis workgroup (4 waves x 16-row MFMA tiles).
constexpr int kRingsPerWG = 64;
// Degrees per analysis workgroup (2 parities x 2 waves x 16 rows x degree tiles per wave):
// two tiles per wave in the split modes, where one staged (and split) chunk of Fourier rows
// then serves 128 degrees, one in FP32 mode, where staging is cheap and registers are not.
constexpr int ana_tiles_per_wave(int mode) { return mode == 0 ? 1 : 2; }
constexpr int ana_degs_per_wg(int mode) { return 64 * ana_tiles_per_wave(mode); }
constexpr int kDegPerWGMax = 128;
// Contraction chunk staged in LDS: 32 degrees-per-parity (synthesis) or 32 rings
// (analysis), one MFMA K-step of the split modes.
constexpr int kChunk = 32;
constexpr int kMaxFactors = 24;
constexpr int kFFTThreads = 256;
// Largest ring the FFT kernel takes, and columns per FFT workgroup.
constexpr int kFFTMaxRing = 1024;
constexpr int kFFTCols = 4;
// A first radix of at least kMfmaMinPrime runs as the MFMA pass; the factorisation puts
// every prime from kMfmaMinRadix up first (the largest one), and 11 or 13 first only when
// the tuned policy asks for it.  Later passes run as register butterflies, which exist for
// 2, 3, 4, 5, 7, 8, 11, 13.  Rings of nlon = 4 m, m <= 256 (the supported grids) have at
// most one prime factor >= 17, which always takes the MFMA pass.
constexpr int kMfmaMinPrime = 11;
constexpr int kMfmaMinRadix = 17;
constexpr int kMaxRegisterRadix = 13;
// MFMA prime pass of the FFT: largest radix (its tables are indexed by the radix) and the
// row tiles (16 output frequencies each) one wave holds for its column tile in a round.
constexpr int kMaxMfmaRadix = 255;
constexpr int kPrimeRowTiles = 4;
// Row tiles per wave of the MFMA prime pass for radix r on M = N / r butterflies per column
// and ncol columns, maximised over the rounds (host check of kPrimeRowTiles).
inline int prime_rows_per_wave(int r, int M, int ncol) {
    const int Mt = ((r - 1) / 2 + 16) / 16, Nt = (2 * ncol * M + 15) / 16, NW = kFFTThreads / 64;
    int worst = 0;
    for (int n0 = 0; n0 < Nt; n0 += NW) {
        const int nc = Nt - n0 < NW ? Nt - n0 : NW, G = NW / nc;
        const int rows = (Mt + G - 1) / G;
        worst = rows > worst ? rows : worst;
    }
    return worst;
}

enum Mode { kFP32 = 0, kFP16x3 = 1, kBF16x3 = 2 };
// 16-bit parts per operand value: scaled FP16 hi + lo (three products), BF16 hi + mid + lo
// (six products: every term down to 2^-16 of the leading one).
constexpr int split_parts(int mode) { return mode == kBF16x3 ? 3 : 2; }

// One workgroup's work item.  Legendre stages: order m, ring group (synthesis) or degree
// group (analysis), column group cg.  FFT stages: m = ring, blk = column group.  call
// indexes the launch's call arguments (always 0 for a single transform).
struct Tile {
    int call;  // batch Legendre tiles: < 0 for merged column group -1 - call (see ColRef)
    int m;     // < 0: padding, the workgroup exits at once
    int blk;
    int cg;
};

// A column of a merged Legendre column group: batch items on one plan (and, for the
// synthesis, one unscale_coslat flag) share column groups of kColsComplex columns, so one
// workgroup (order m, ring or degree block) reads Λ once for all of them.  Group q's columns
// are entries q kColsComplex .. + kColsComplex - 1 of the batch's column map, which batch
// launches pass as single.in; call < 0 marks an empty column.
struct ColRef {
    int call;
    int col;
};

// Workgroups go to the 8 XCDs (each with its own L2) round-robin by linear block index.
// Launch orders place workgroups that read the same data (the tiles of one order m; the
// two FFT column groups that share 128-byte Fourier lines) on one XCD, consecutively.
constexpr int kXCDs = 8;

// Arguments of one transform call inside a launch.
struct CallArgs {
    int plan;             // index into the launch's plan array
    int K;                // columns
    int unscale;          // synthesis: divide by cos(lat)
    int paired;           // single-call FFT synthesis: a ring's column groups on one XCD
    const void *in;       // synthesis: coefficients; analysis: grid
    void *out;            // synthesis: grid; analysis: coefficients
    long long ld_in;      // column strides (complex elements for spectra, floats for grids)
    long long ld_out;
    void *fbuf;           // Fourier scratch of this call
    // Adjoint (pullback) calls: 1 = the FFT analysis scales its Fourier rows by
    // s_m (1/cos lat) / ΔΩ_j, 2 = the FFT synthesis scales its input rows by ΔΩ_j / s_m, with
    // s_m = 2 for interior bins and 1 for DC and Nyquist; accumulate: the last stage adds
    // into its output instead of overwriting it.
    int adj;
    int accumulate;
    // Fused spectral Laplacian (spectral_gradients.jl:385-466): lap 0 none, 1 ∇² (degree
    // factor e_l = -l(l+1)), 2 ∇⁻² (1/e_l, 0 at l = 0).  The Legendre synthesis multiplies
    // its input degrees by lap_sign e_l lap_scale; the Legendre analysis multiplies its
    // output by lap_sign e_l, adds the old output if accumulate, then scales by lap_scale
    // (the reference's order).  Zero-initialised calls (lap = 0) are plain transforms.
    int lap;
    float lap_sign;
    float lap_scale;
};

enum Stage { kLegendreSynthesis = 0, kFFTSynthesis = 1, kFFTAnalysis = 2, kLegendreAnalysis = 3 };

// Fourier-stage kernels: the mixed-radix LDS FFT, or the ring DFT on MFMA (a radix-4 pass
// on the vector units and a p-point DFT as two GEMMs, FP32 or scaled FP16 hi/lo products),
// with 4 or 8 columns per workgroup.
enum FourierKind { kFourierRadix = 0, kFourierDft32 = 1, kFourierDft16 = 2 };
constexpr int kDftChunk32 = 16;  // contraction values per FP32 fragment load (4 K-steps of 4)
constexpr int kDftChunk16 = 32;  // contraction values per FP16 MFMA K-step
constexpr float kDftAScale = 16384.f;  // FP16 tables hold 2^14 cos, 2^14 sin
// Rows of the DFT GEMM an analysis of ring (p, mt) needs: r <= min(p/2, mt/4 + 1).
__host__ __device__ inline int dft_rows_analysis(int p, int mt) {
    const int a = p / 2, b = mt / 4 + 1;
    return (a < b ? a : b) + 1;
}
// Contraction length of a synthesis: t <= min((p-1)/2, (mt+3)/4) (orders above mt are zero).
__host__ __device__ inline int dft_k_synthesis(int p, int mt) {
    const int a = (p - 1) / 2, b = (mt + 3) / 4;
    return a < b ? a : b;
}
// Padded contraction length and LDS row strides (elements) of the staged B operands.
__host__ __device__ inline int dft_kpad(int K, bool f16) {
    const int c = f16 ? kDftChunk16 : kDftChunk32;
    return (K + c - 1) / c * c;
}
__host__ __device__ inline int dft_bstride(int kpad, bool f16) { return kpad + (f16 ? 8 : 4); }
// Dynamic LDS bytes of one DFT workgroup (cg columns) for a ring (p, mt); the launch takes
// the maximum over the rings.  Analysis: S, D operands then the staged outputs F (aliased);
// synthesis: staged Fourier rows, S, D operands, then the outputs Y (aliased over both).
__host__ __device__ inline int dft_smem_bytes(bool synthesis, bool f16, int cg, int p, int mt) {
    const int h = (p - 1) / 2;
    const int K = synthesis ? dft_k_synthesis(p, mt) : h;
    const int kp = dft_kpad(K < 1 ? 1 : K, f16);
    const int bs = dft_bstride(kp, f16);
    const int cols = 8 * cg;
    const int bbytes = 2 * cols * bs * (f16 ? 4 : 4);  // S and D: float, or FP16 hi + lo
    const int aux = cg * 4 * 2 * 8 + cg * 4 * 4;        // X0, Xh per (column, m2), scales
    if (!synthesis) {
        const int fbytes = cg * (2 * mt + 2) * 8;
        return (bbytes > fbytes ? bbytes : fbytes) + aux;
    }
    const int zbytes = cg * (mt + 1) * 16;
    const int ybytes = cg * p * 4 * 8;
    const int a = zbytes + bbytes;
    return (a > ybytes ? a : ybytes) + aux;
}

// One pass of a ring FFT (radix R, l = product of the earlier radices, M = N / (l R)) with
// the multipliers of FastDiv for M and R precomputed, so the kernels divide by neither.
struct FFTPass {
    int R, l, M;
    unsigned mulM, mulR;
    int pad[3];
};
// FastDiv multiplier of d: x / d == umulhi(x, mul) for x, d < 2^16.
inline unsigned fastdiv_mul(int d) { return d > 1 ? 0xFFFFFFFFu / (unsigned)d + 1u : 0u; }

// One FFT workgroup rank's ring (launch order, heaviest first, per factorisation policy):
// everything the FFT kernels read per ring in one record, so that a workgroup's first loads
// depend on no other load.  prime, M0, mulM0, mulR0: the MFMA first pass (radix, butterflies
// per column, FastDiv multipliers of M0 and prime), prime = 0 without one.
struct FFTRing {
    int j, N, mt, mte;
    int nf, prime, M0;
    unsigned mulM0, mulR0;
    int pad[3];
    long long tw_off;     // first twiddle of the ring
    long long fboff;      // first Fourier-buffer row
    long long ring_n;     // first grid point of the northern ring
    long long ring_s;     // first grid point of the southern ring
    long long prime_off;  // the MFMA first pass's table (DevPlan::prime_tab, 8-element vectors)
    long long passes;     // first pass descriptor (index into DevPlan::passes)
};
static_assert(sizeof(FFTRing) == 96, "FFTRing layout");

// Device view of a plan (kernels read it through a device pointer).
struct DevPlan {
    int n;            // rings per hemisphere
    int lmax, mmax;   // 1-based
    int nharm;
    long long npoints;
    int kg;           // columns per FFT workgroup
    int fft_smem;     // dynamic LDS bytes of the FFT kernel

    // per northern ring j
    const int *nlon;
    const int *mt;            // 0-based last order kept (Legendre shortcut)
    const int *mte;           // max of mt over rings 0..j: the FFT analysis writes rows
                              // mt < m <= mte as zeros (Legendre stage rings of order m)
    const long long *fboff;   // first Fourier-buffer row of ring j (sum of nfreq before j)
    const long long *ring_n;  // first grid point of ring j
    const long long *ring_s;  // first grid point of the mirrored southern ring
    const float *coslat_inv;
    const float *domega;      // solid angle ΔΩ_j of a grid point on ring j
    const int *ring_order;    // FFT workgroup -> ring, most expensive first
    const FFTRing *rings;     // FFT workgroup rank -> its ring's record (same order)

    // per order m
    const int *j0;               // first ring with mt >= m (rings j0..n-1 contribute)
    const long long *lm_off;     // packed offset of column m
    const long long *syn_off;    // offset of m's block in the synthesis Λ arrays
    const long long *ana_off;    // offset of m's block in the analysis Λ arrays
    const long long *syn_soff;   // offset of m's row scales (synthesis, per ring)
    const long long *ana_soff;   // offset of m's row scales (analysis, per degree)

    // Λ in MFMA-fragment order.  FP32 mode: float arrays.  Split modes: arrays of 16-bit
    // parts, v = hi + lo (scaled FP16) or v = hi + mid + lo (BF16, three terms), and
    // inverse row scales (FP16 only).
    const float *syn_f32;
    const void *syn_hi;
    const void *syn_lo;
    const void *syn_mid;
    const float *syn_sinv;
    const float *ana_f32;
    const void *ana_hi;
    const void *ana_lo;
    const void *ana_mid;
    const float *ana_sinv;

    // FFT tables
    const float *twiddle;  // interleaved complex exp(-2 pi i k / N) for each ring
    const long long *tw_off;
    const FFTPass *passes;  // kMaxFactors passes per ring, in order
    const int *nfac;
    // First-pass prime radices r (MFMA pass): per r, cos / sin(2 pi s t / r) for rows
    // s = 0..h (tiles of 16) and t = 1..h (chunks of 32), h = (r-1)/2, as three BF16 parts of
    // the FP64 values in MFMA fragment order: per (row tile, chunk, lane) six vectors of 8
    // (cos hi, mid, lo, sin hi, mid, lo).  prime_off[r] (r < 256): r's first vector.
    const unsigned short *prime_tab;
    const long long *prime_off;
    // Ring DFT on MFMA (rings of N = 4p points): per ring, the cos / sin matrices
    // cos(2 pi r t / p), sin(2 pi r t / p), rows r = 0..p/2 (tiles of 16), columns
    // t = 1..(p-1)/2, in MFMA fragment order: FP32 (dft32, chunks of 16 t, a float4 of cos
    // then one of sin per lane) and, in the split modes, FP16 hi / lo of 2^14 times the value
    // (dft16, chunks of 32 t, per lane 8 cos hi, 8 cos lo, 8 sin hi, 8 sin lo).
    const float *dft32;
    const long long *dft32_off;
    const unsigned short *dft16;
    const long long *dft16_off;

    // UV_from_vordiv factors (per packed lm): x = -m/(l(l+1)), u1 = eps_lm / l,
    // u2 = eps_{l+1,m} / (l+1) (0 on the last row)
    const float *uv_x;
    const float *uv_1;
    const float *uv_2;
    const int *l_of;       // 0-based degree of packed index
    // gradient operator factors per packed lm (gradient_arrays.jl:27-100):
    // gy = (grad_y_vordiv1, grad_y_vordiv2, grad_y1, grad_y2), gx = (grad_x_vordiv, m)
    const float4 *gy;
    const float2 *gx;
};

// Records the message sht_last_error() returns on this thread.
void set_last_error(const std::string &msg);

// A plain transform call (SHT_DIR_SYNTHESIS or SHT_DIR_ANALYSIS, the conventions of
// sht_synthesis / sht_analysis) whose kernels use `scratch` (at least plan_scratch_bytes(plan,
// K) bytes of device memory) instead of the stream's scratch of the plan.
int transform_with_scratch(sht_plan *plan, int direction, const void *in, int64_t ld_in, void *out, int64_t ld_out,
                           int32_t K, int32_t unscale_coslat, void *scratch, void *stream);
size_t plan_scratch_bytes(const sht_plan *plan, int K);
int plan_device(const sht_plan *plan);
// Looks up the address of the static scratch array (sht_static.hip) on `device`, which
// must be current; once per device, later calls return at once.
int static_scratch_setup(int device);

// Kernel entry points.  Every stage kernel takes
//   (const DevPlan *plans, CallArgs single, const CallArgs *calls, const Tile *tiles)
// on a 1-D grid.  batch = false: the single call's arguments by value; the Legendre
// kernels read their tile from tiles[blockIdx.x], the FFT kernels derive ring and column
// group from blockIdx.x.  batch = true: tile, call and column group from tiles.
const void *legendre_kernel(int stage, int mode, bool batch);
// accumulate: the single-call synthesis that adds into its output (pullbacks); batches
// never accumulate.
const void *fft_kernel(int stage, bool batch, bool accumulate = false);
// Ring DFT on MFMA (sht_dft.hip): FP16 hi/lo or FP32 products, cg = 4 or 8 columns per
// workgroup; single calls launch n * ceil(K / cg) blocks, dynamic LDS dft_smem_bytes.
const void *dft_kernel(int stage, bool f16, int cg, bool batch, bool accumulate = false);
// Spectral gradient operators (one kernel, op code; flags: kOp* bits).
enum SpectralOp { kOpDivergence = 0, kOpCurl = 1, kOpLaplace = 2, kOpGradient = 3 };
constexpr int kOpFlip = 1, kOpAdd = 2, kOpInverse = 4;
int launch_spectral_op(const DevPlan *P, int nharm, int op, const void *a, const void *b, void *o1, void *o2,
                       long long ld, int K, int flags, float rscale, void *stream);
// field[k][p] *= scale[p] for p < npoints, k < K (column stride ld)
int launch_scale_points(const float *scale, void *field, long long npoints, long long ld, int K, void *stream);

// Dense transform GEMM on MFMA (split-K): launch shape of C = A B with A M x N, K columns.
struct DenseSplit {
    long long rt;  // row tiles of 16
    long long s4;  // fragment groups of 16 contraction values
    int cg;        // column groups of 64
    int s;         // splits of the contraction (1: C written directly)
    int s4_split;  // fragment groups per split
};
DenseSplit dense_split(long long M, long long N, int K);
// Position of A[r][n] (M x N) in the fragment order the GEMM reads (an array of
// 16 ceil(M/16) x 16 ceil(N/16) floats, zero where r >= M or n >= N).
long long dense_fragment_index(long long r, long long n, long long S4);
// C = A B (optionally rows times row_scale), A in fragment order; work holds
// dense_split(M, N, K).s * K * M floats when that split count exceeds 1.
int launch_dense_gemm(const float *Af, long long M, long long N, const float *B, long long ldb, int K, float *C,
                      long long ldc, const float *row_scale, float *work, void *stream);
int launch_uv_from_vordiv(const DevPlan *P, int nharm, const void *vor, const void *div, void *U,
                          void *V, long long ld, int K, float radius, void *stream);

}  // namespace sht
