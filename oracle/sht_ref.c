/* CPU restatement of SpeedyTransforms' transform! in FP32 and FP64.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): built into oracle/liboracle.so and loaded by
 * tests/ and by bench.py's cpu_baseline leg, never by the product library.
 *
 * The Legendre loops follow SpeedyTransforms/src/legendre.jl:2-33,38-101 (inverse,
 * even/odd fused dot products with muladd) and :104-126,131-198 (forward, outer-product
 * accumulation over rings in ascending order), with the linear shortcut of
 * legendre_shortcuts.jl:21-25 given as mtrunc[j].  The Fourier stage restates the per-ring
 * unnormalised rfft / brfft of fourier.jl:27-77 (FFTW semantics: Im of DC and Nyquist
 * ignored by brfft) with a mixed-radix FFT written here, transforming the northern and
 * the mirrored southern ring together as one complex sequence.  Tables (Λ in [ring][lm]
 * order, solid angles, 1/cos lat, ring offsets) come from the caller.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <omp.h>

typedef struct {
    int32_t nlat_half;
    int32_t lmax, mmax;          /* 1-based */
    int32_t nharm;
    int64_t npoints;
    const int32_t *nlons;        /* [nlat_half] points on northern ring j */
    const int64_t *ring_start;   /* [2 nlat_half] first point of ring, north to south */
    const int32_t *mtrunc;       /* [nlat_half] 0-based max order per ring */
    const int64_t *offsets;      /* [mmax] packed start of column m */
    const double *coslat_inv;    /* [nlat_half] */
    const double *solid_angles;  /* [nlat_half] */
} oref_geom;

#define REAL float
#define SFX _f32
#define FFTPLAN_T fftplan_f32
#define FFT_INIT fft_plan_init_f32
#define FFT_EXEC fft_exec_f32
#define FFT_FREE fft_plan_free_f32
#include "sht_ref_impl.h"
#undef REAL
#undef SFX
#undef FFTPLAN_T
#undef FFT_INIT
#undef FFT_EXEC
#undef FFT_FREE

#define REAL double
#define SFX _f64
#define FFTPLAN_T fftplan_f64
#define FFT_INIT fft_plan_init_f64
#define FFT_EXEC fft_exec_f64
#define FFT_FREE fft_plan_free_f64
#include "sht_ref_impl.h"

/* plain complex DFT helpers for the FFT self-test */
int oref_fft_f64(int n, int sign, const double *in, double *out) {
    fftplan_f64 pl;
    fft_plan_init_f64(&pl, n, sign);
    cpx_f64 *work = (cpx_f64 *)malloc(sizeof(cpx_f64) * (size_t)n);
    fft_exec_f64(&pl, (cpx_f64 *)out, (const cpx_f64 *)in, work);
    free(work);
    fft_plan_free_f64(&pl);
    return 0;
}

int oref_max_threads(void) { return omp_get_max_threads(); }
