// Internal definitions shared by the host plan code and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace sht {

// Column group: 32 complex columns = 64 real MFMA columns per Legendre workgroup.
constexpr int kColsComplex = 32;
constexpr int kColsReal = 64;
// Rings per synthesis workgroup (4 waves x 16-row MFMA tiles).
constexpr int kRingsPerWG = 64;
// Degrees per analysis workgroup (2 parities x 2 waves x 16 rows).
constexpr int kDegPerWG = 64;
// Contraction chunk staged in LDS: 32 degrees-per-parity (synthesis) or 32 rings
// (analysis), one MFMA K-step of the split modes.
constexpr int kChunk = 32;
constexpr int kMaxFactors = 24;
constexpr int kFFTThreads = 256;

enum Mode { kFP32 = 0, kFP16x3 = 1, kBF16x3 = 2 };

struct LegTile {
    int m;    // order (0-based)
    int blk;  // ring group (synthesis) or degree group (analysis)
};

// Device view of a plan, passed by value to every kernel.
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
    const long long *fboff;   // first Fourier-buffer row of ring j (sum of nfreq before j)
    const long long *ring_n;  // first grid point of ring j
    const long long *ring_s;  // first grid point of the mirrored southern ring
    const float *coslat_inv;
    const int *ring_order;    // FFT workgroup -> ring, most expensive first

    // per order m
    const int *j0;               // first ring with mt >= m (rings j0..n-1 contribute)
    const long long *lm_off;     // packed offset of column m
    const long long *syn_off;    // offset of m's block in the synthesis Λ arrays
    const long long *ana_off;    // offset of m's block in the analysis Λ arrays
    const long long *syn_soff;   // offset of m's row scales (synthesis, per ring)
    const long long *ana_soff;   // offset of m's row scales (analysis, per degree)

    // Λ in MFMA-fragment order.  FP32 mode: float arrays.  Split modes: hi and lo
    // arrays of 16-bit values (FP16 or BF16) and inverse row scales (FP16 only).
    const float *syn_f32;
    const void *syn_hi;
    const void *syn_lo;
    const float *syn_sinv;
    const float *ana_f32;
    const void *ana_hi;
    const void *ana_lo;
    const float *ana_sinv;

    // FFT tables
    const float *twiddle;  // interleaved complex exp(-2 pi i k / N) for each ring
    const long long *tw_off;
    const int *fac;        // kMaxFactors radices per ring
    const int *nfac;

    // UV_from_vordiv factors (per packed lm): x = -m/(l(l+1)), u1 = eps_lm / l,
    // u2 = eps_{l+1,m} / (l+1) (0 on the last row)
    const float *uv_x;
    const float *uv_1;
    const float *uv_2;
    const int *l_of;       // 0-based degree of packed index
};

// Kernel launchers (sht_kernels.hip).
int launch_legendre_synthesis(const DevPlan &P, int mode, const LegTile *tiles, int ntiles,
                              const void *coeffs, long long ld_spec, void *fbuf, int K,
                              int unscale, void *stream);
int launch_legendre_analysis(const DevPlan &P, int mode, const LegTile *tiles, int ntiles,
                             const void *fbuf, void *coeffs, long long ld_spec, int K,
                             void *stream);
int launch_fft_synthesis(const DevPlan &P, const void *fbuf, void *field, long long ld_grid,
                         int K, void *stream);
int launch_fft_analysis(const DevPlan &P, const void *field, long long ld_grid, void *fbuf,
                        int K, void *stream);
int launch_uv_from_vordiv(const DevPlan &P, const void *vor, const void *div, void *U,
                          void *V, long long ld, int K, float radius, void *stream);

}  // namespace sht
