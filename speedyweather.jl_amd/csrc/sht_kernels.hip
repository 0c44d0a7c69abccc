// gfx950 kernels of the spherical harmonic transform.
//
//  * Legendre stage: for every order m a dense contraction between the precomputed
//    Legendre polynomials Λ_m (degrees x rings) and the K spectral / Fourier columns,
//    split into even and odd degrees l-m so one contraction serves both hemispheres
//    (the even/odd symmetry of SpeedyTransforms/src/legendre.jl:2-33,104-126).  Runs on
//    MFMA: v_mfma_f32_16x16x4_f32 in FP32 mode; v_mfma_f32_16x16x32_{f16,bf16} with a
//    hi/lo split of both operands (3 products, FP32 accumulation) in the split modes.
//  * Fourier stage: one complex FFT of length nlon per ring carrying the northern ring
//    in the real and the mirrored southern ring in the imaginary part (the per-ring
//    rfft / brfft pairs of SpeedyTransforms/src/fourier.jl:27-77), a mixed-radix
//    Stockham FFT in LDS.
//
// Intermediate Fourier buffer (owned by the plan): row (fboff[j] + m) holds, for every
// column k, a float4 (N.re, N.im, S.re, S.im).  Synthesis writes the northern/southern
// Legendre sums g_N, g_S there (legendre.jl:84-91); analysis writes the ΔΩ-free
// even/odd combinations f_N + f_S, f_N - f_S (legendre.jl:176-183, ΔΩ folded into Λ).
#include <hip/hip_runtime.h>
#include <math.h>
#include "sht_internal.h"

namespace sht {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;
typedef u16 u16x8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ MFMA helpers

__device__ inline f4 mfma_f32(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int MODE>
struct Split;

template <>
struct Split<kFP16x3> {
    typedef _Float16 T;
    typedef h8 V;
    __device__ static inline f4 mma(V a, V b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Split<kBF16x3> {
    typedef __bf16 T;
    typedef b8 V;
    __device__ static inline f4 mma(V a, V b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// Power-of-two scale that maps max|x| into [2^14, 2^15) (FP16 hi/lo), with its inverse.
__device__ inline void fp16_scale(float amax, float &s, float &sinv) {
    if (!(amax > 0.f) || !isfinite(amax)) {
        s = 1.f;
        sinv = 1.f;
        return;
    }
    int e;
    frexpf(amax, &e);  // amax = f 2^e, f in [0.5, 1)
    int sh = 15 - e;
    sh = sh > 120 ? 120 : (sh < -120 ? -120 : sh);
    s = ldexpf(1.f, sh);
    sinv = ldexpf(1.f, -sh);
}

// ============================================================ Legendre synthesis
//
// Workgroup = (order m, group of 64 rings, 32 complex columns).  Wave w owns ring tile
// 4*blk+w (16 rings) and all 64 real columns, for both parities: 8 accumulators.
// Degrees l = m + 2l' + p are streamed in chunks of 32 l' per parity through LDS.

template <int MODE>
__global__ __launch_bounds__(256) void legendre_synthesis_kernel(
    DevPlan P, const LegTile *__restrict__ tiles, const float2 *__restrict__ coeffs,
    long long ld_spec, float4 *__restrict__ fbuf, int K, int unscale) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const LegTile T = tiles[blockIdx.x];
    const int m = T.m, cg = blockIdx.y;
    const int n = P.n;
    const int j0 = P.j0[m];
    const int J = n - j0;
    const int L = P.lmax - m;
    const int Le = (L + 1) >> 1;
    const int nch = (Le + kChunk - 1) / kChunk;
    const int Jt = (J + 15) >> 4;
    const int tile = T.blk * 4 + w;
    const bool active = tile < Jt;
    const long long off = P.lm_off[m];
    const int kbase = cg * kColsComplex;
    const int row = lane & 15, grp = lane >> 4;

    __shared__ float Bf[2][kChunk][kColsReal];  // staged coefficients [parity][l'][col]
    constexpr int BP = kChunk + 8;               // padded row (16-bit elements)
    __shared__ alignas(16) u16 Bh[2][kColsReal][BP];  // split modes: hi [parity][col][l']
    __shared__ alignas(16) u16 Bl[2][kColsReal][BP];
    __shared__ float Bsinv[2][kColsReal];

    f4 acc[2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[p][cb] = f4{0.f, 0.f, 0.f, 0.f};

    for (int c = 0; c < nch; ++c) {
        __syncthreads();
        // stage 64 consecutive degrees x 32 complex columns
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int idx = it * 256 + tid;
            const int kk = idx >> 6, o = idx & 63;
            const int og = 64 * c + o, kg = kbase + kk;
            float2 v = make_float2(0.f, 0.f);
            if (og < L && kg < K) v = coeffs[(long long)kg * ld_spec + off + og];
            Bf[o & 1][o >> 1][2 * kk] = v.x;
            Bf[o & 1][o >> 1][2 * kk + 1] = v.y;
        }
        __syncthreads();
        if constexpr (MODE == kFP32) {
            if (active) {
                const float *A = P.syn_f32 + P.syn_off[m];
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const float *Ap = A + ((long long)(p * Jt + tile) * (nch * kChunk) + c * kChunk) * 16;
#pragma unroll
                    for (int ks = 0; ks < kChunk / 4; ++ks) {
                        const float a = Ap[(ks * 4 + grp) * 16 + row];
#pragma unroll
                        for (int cb = 0; cb < 4; ++cb) {
                            const float b = Bf[p][ks * 4 + grp][cb * 16 + row];
                            acc[p][cb] = mfma_f32(a, b, acc[p][cb]);
                        }
                    }
                }
            }
        } else {
            typedef typename Split<MODE>::T T16;
            typedef typename Split<MODE>::V V8;
            if (tid < 2 * kColsReal) {
                const int p = tid >> 6, col = tid & 63;
                float s = 1.f, sinv = 1.f;
                if constexpr (MODE == kFP16x3) {
                    float amax = 0.f;
                    for (int k2 = 0; k2 < kChunk; ++k2) amax = fmaxf(amax, fabsf(Bf[p][k2][col]));
                    fp16_scale(amax, s, sinv);
                }
                Bsinv[p][col] = sinv;
                for (int k2 = 0; k2 < kChunk; ++k2) {
                    const float x = Bf[p][k2][col] * s;
                    const T16 hi = (T16)x;
                    const T16 lo = (T16)(x - (float)hi);
                    Bh[p][col][k2] = __builtin_bit_cast(u16, hi);
                    Bl[p][col][k2] = __builtin_bit_cast(u16, lo);
                }
            }
            __syncthreads();
            if (active) {
                const V8 *Ah = reinterpret_cast<const V8 *>(static_cast<const u16 *>(P.syn_hi) + P.syn_off[m]);
                const V8 *Al = reinterpret_cast<const V8 *>(static_cast<const u16 *>(P.syn_lo) + P.syn_off[m]);
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const long long fi = ((long long)(p * Jt + tile) * nch + c) * 64 + lane;
                    const V8 ah = Ah[fi], al = Al[fi];
#pragma unroll
                    for (int cb = 0; cb < 4; ++cb) {
                        const int col = cb * 16 + row;
                        const V8 bh = __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(&Bh[p][col][8 * grp]));
                        const V8 bl = __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(&Bl[p][col][8 * grp]));
                        if constexpr (MODE == kFP16x3) {
                            f4 t = Split<MODE>::mma(al, bh, f4{0.f, 0.f, 0.f, 0.f});
                            t = Split<MODE>::mma(ah, bl, t);
                            t = Split<MODE>::mma(ah, bh, t);
                            const float sb = Bsinv[p][col];
                            acc[p][cb] += t * sb;
                        } else {
                            acc[p][cb] = Split<MODE>::mma(al, bh, acc[p][cb]);
                            acc[p][cb] = Split<MODE>::mma(ah, bl, acc[p][cb]);
                            acc[p][cb] = Split<MODE>::mma(ah, bh, acc[p][cb]);
                        }
                    }
                }
            }
        }
    }

    if (!active) return;  // no further barriers below
    // epilogue: N = E + O, S = E - O (legendre.jl:84-91), optional 1/cos(lat)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int col = cb * 16 + row;
        const int kk = col >> 1, reim = col & 1;
        const int k = kbase + kk;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 4 * grp + i;
            const int j = j0 + 16 * tile + r;
            float e = acc[0][cb][i], o = acc[1][cb][i];
            if constexpr (MODE == kFP16x3) {
                const float sa = P.syn_sinv[P.syn_soff[m] + 16 * tile + r];
                e *= sa;
                o *= sa;
            }
            float nv = e + o, sv = e - o;
            if (unscale && j < n) {
                const float ci = P.coslat_inv[j];
                nv *= ci;
                sv *= ci;
            }
            const float n_other = __shfl_xor(nv, 1);
            const float s_other = __shfl_xor(sv, 1);
            if (j < n && k < K) {
                float *dst = reinterpret_cast<float *>(fbuf + (P.fboff[j] + m) * (long long)K + k);
                if (reim == 0) {
                    *reinterpret_cast<float2 *>(dst) = make_float2(nv, n_other);
                } else {
                    *reinterpret_cast<float2 *>(dst + 2) = make_float2(s_other, sv);
                }
            }
        }
    }
}

// ============================================================== Legendre analysis
//
// Workgroup = (order m, group of 64 degrees, 32 complex columns).  Wave w owns parity
// p = w&1 and degree tile lt = 2*blk + (w>>1) (16 rows of l' with l = m + 2l' + p) for
// all 64 real columns.  Rings are streamed in chunks of 32 through LDS.  Every output
// coefficient is written exactly once (no atomics, deterministic).

template <int MODE>
__global__ __launch_bounds__(256) void legendre_analysis_kernel(
    DevPlan P, const LegTile *__restrict__ tiles, const float4 *__restrict__ fbuf,
    float2 *__restrict__ coeffs, long long ld_spec, int K) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const LegTile T = tiles[blockIdx.x];
    const int m = T.m, cg = blockIdx.y;
    const int n = P.n;
    const int j0 = P.j0[m];
    const int J = n - j0;
    const int L = P.lmax - m;
    const int NG = (L + kDegPerWG - 1) / kDegPerWG;
    const int nt = 2 * NG;
    const int JC = (J + kChunk - 1) / kChunk;
    const int p = w & 1;
    const int lt = 2 * T.blk + (w >> 1);
    const long long off = P.lm_off[m];
    const int kbase = cg * kColsComplex;
    const int row = lane & 15, grp = lane >> 4;

    constexpr int BP = kChunk + 8;
    __shared__ union {
        float Bf[2][kChunk][kColsReal];          // staged Fourier data [parity][ring][col]
        float2 Out[kColsComplex][kDegPerWG + 1]; // output tile [k][l - m - 64 blk]
    } sm;
    __shared__ alignas(16) u16 Bh[2][kColsReal][BP];
    __shared__ alignas(16) u16 Bl[2][kColsReal][BP];
    __shared__ float Bsinv[2][kColsReal];

    f4 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = f4{0.f, 0.f, 0.f, 0.f};

    for (int c = 0; c < JC; ++c) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int idx = it * 256 + tid;
            const int jj = idx >> 5, kk = idx & 31;
            const int j = j0 + kChunk * c + jj, kg = kbase + kk;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j < n && kg < K) v = fbuf[(P.fboff[j] + m) * (long long)K + kg];
            sm.Bf[0][jj][2 * kk] = v.x;
            sm.Bf[0][jj][2 * kk + 1] = v.y;
            sm.Bf[1][jj][2 * kk] = v.z;
            sm.Bf[1][jj][2 * kk + 1] = v.w;
        }
        __syncthreads();
        if constexpr (MODE == kFP32) {
            const float *A = P.ana_f32 + P.ana_off[m] +
                             ((long long)(p * nt + lt) * (JC * kChunk) + c * kChunk) * 16;
#pragma unroll
            for (int ks = 0; ks < kChunk / 4; ++ks) {
                const float a = A[(ks * 4 + grp) * 16 + row];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    const float b = sm.Bf[p][ks * 4 + grp][cb * 16 + row];
                    acc[cb] = mfma_f32(a, b, acc[cb]);
                }
            }
        } else {
            typedef typename Split<MODE>::T T16;
            typedef typename Split<MODE>::V V8;
            if (tid < 2 * kColsReal) {
                const int pp = tid >> 6, col = tid & 63;
                float s = 1.f, sinv = 1.f;
                if constexpr (MODE == kFP16x3) {
                    float amax = 0.f;
                    for (int k2 = 0; k2 < kChunk; ++k2) amax = fmaxf(amax, fabsf(sm.Bf[pp][k2][col]));
                    fp16_scale(amax, s, sinv);
                }
                Bsinv[pp][col] = sinv;
                for (int k2 = 0; k2 < kChunk; ++k2) {
                    const float x = sm.Bf[pp][k2][col] * s;
                    const T16 hi = (T16)x;
                    const T16 lo = (T16)(x - (float)hi);
                    Bh[pp][col][k2] = __builtin_bit_cast(u16, hi);
                    Bl[pp][col][k2] = __builtin_bit_cast(u16, lo);
                }
            }
            __syncthreads();
            const V8 *Ah = reinterpret_cast<const V8 *>(static_cast<const u16 *>(P.ana_hi) + P.ana_off[m]);
            const V8 *Al = reinterpret_cast<const V8 *>(static_cast<const u16 *>(P.ana_lo) + P.ana_off[m]);
            const long long fi = ((long long)(p * nt + lt) * JC + c) * 64 + lane;
            const V8 ah = Ah[fi], al = Al[fi];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const int col = cb * 16 + row;
                const V8 bh = __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(&Bh[p][col][8 * grp]));
                const V8 bl = __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(&Bl[p][col][8 * grp]));
                if constexpr (MODE == kFP16x3) {
                    f4 t = Split<MODE>::mma(al, bh, f4{0.f, 0.f, 0.f, 0.f});
                    t = Split<MODE>::mma(ah, bl, t);
                    t = Split<MODE>::mma(ah, bh, t);
                    acc[cb] += t * Bsinv[p][col];
                } else {
                    acc[cb] = Split<MODE>::mma(al, bh, acc[cb]);
                    acc[cb] = Split<MODE>::mma(ah, bl, acc[cb]);
                    acc[cb] = Split<MODE>::mma(ah, bh, acc[cb]);
                }
            }
        }
    }
    __syncthreads();
    // stage the 64-degree x 32-column output tile, then write l-contiguous columns
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int col = cb * 16 + row;
        const int kk = col >> 1, reim = col & 1;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 4 * grp + i;
            const int lq = 16 * (w >> 1) + r;  // l' - 32 blk
            float v = acc[cb][i];
            if constexpr (MODE == kFP16x3) v *= P.ana_sinv[P.ana_soff[m] + (p * nt + lt) * 16 + r];
            float *o = reinterpret_cast<float *>(&sm.Out[kk][2 * lq + p]);
            o[reim] = v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int idx = it * 256 + tid;
        const int kk = idx >> 6, o = idx & 63;
        const int og = kDegPerWG * T.blk + o, kg = kbase + kk;
        if (og < L && kg < K) coeffs[(long long)kg * ld_spec + off + og] = sm.Out[kk][o];
    }
}

// ====================================================================== ring FFT
//
// Stockham autosort, decimation in time.  With N = r_1 r_2 ... r_s, pass p of radix r
// (l = r_1...r_{p-1}, M = N/(l r)) maps, for b in [0,M), u in [0,l):
//   in[b + M(q + r u)], q < r   ->   out[b + M(u + l s)], s < r
//   out = sum_q (w_N^{q u M} in_q) w_r^{q s},   w_N = exp(-+ 2 pi i / N).
// Radix 2, 3, 4, 5 run as register butterflies; any other radix as a plain DFT
// (one output per thread) after an in-place twiddle pass.

__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ inline float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ inline float2 twid(const float2 *tw, int idx) {
    const float2 t = tw[idx];
    return INV ? make_float2(t.x, -t.y) : t;
}

template <bool INV>
__device__ void fft_pass(const float2 *in, float2 *out, const float2 *tw, int N, int r, int l,
                         int ncol, int tid) {
    const int M = N / (l * r);
    const int nb = N / r;  // butterflies per column
    if (r == 2 || r == 3 || r == 4 || r == 5) {
        for (int it = tid; it < ncol * nb; it += kFFTThreads) {
            const int col = it / nb, t = it - col * nb;
            const int b = t % M, u = t / M;
            const float2 *src = in + col * N + b + M * r * u;
            float2 *dst = out + col * N + b + M * u;
            float2 a[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                if (q < r) {
                    float2 v = src[M * q];
                    if (q > 0 && u > 0) v = cmul(v, twid<INV>(tw, q * u * M));
                    a[q] = v;
                }
            }
            const int ls = M * l;  // output stride between s
            if (r == 2) {
                dst[0] = cadd(a[0], a[1]);
                dst[ls] = csub(a[0], a[1]);
            } else if (r == 4) {
                const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
                const float2 s13 = cadd(a[1], a[3]), d13 = mul_mi<INV>(csub(a[1], a[3]));
                dst[0] = cadd(s02, s13);
                dst[ls] = cadd(d02, d13);
                dst[2 * ls] = csub(s02, s13);
                dst[3 * ls] = csub(d02, d13);
            } else if (r == 3) {
                const float c3 = -0.5f, s3 = 0.86602540378443864676f;
                const float2 s = cadd(a[1], a[2]);
                const float2 d = mul_mi<INV>(csub(a[1], a[2]));
                const float2 base = make_float2(a[0].x + c3 * s.x, a[0].y + c3 * s.y);
                dst[0] = cadd(a[0], s);
                dst[ls] = make_float2(base.x + s3 * d.x, base.y + s3 * d.y);
                dst[2 * ls] = make_float2(base.x - s3 * d.x, base.y - s3 * d.y);
            } else {  // r == 5
                const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
                const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
                const float2 s14 = cadd(a[1], a[4]), d14 = mul_mi<INV>(csub(a[1], a[4]));
                const float2 s23 = cadd(a[2], a[3]), d23 = mul_mi<INV>(csub(a[2], a[3]));
                dst[0] = cadd(a[0], cadd(s14, s23));
                const float2 b1 = make_float2(a[0].x + c1 * s14.x + c2 * s23.x, a[0].y + c1 * s14.y + c2 * s23.y);
                const float2 b2 = make_float2(a[0].x + c2 * s14.x + c1 * s23.x, a[0].y + c2 * s14.y + c1 * s23.y);
                const float2 e1 = make_float2(s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y);
                const float2 e2 = make_float2(s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y);
                dst[ls] = cadd(b1, e1);
                dst[4 * ls] = csub(b1, e1);
                dst[2 * ls] = cadd(b2, e2);
                dst[3 * ls] = csub(b2, e2);
            }
        }
    } else {
        // generic radix: twiddle in place (only where u > 0), then one output per thread
        float2 *inw = const_cast<float2 *>(in);
        if (l > 1) {
            for (int it = tid; it < ncol * N; it += kFFTThreads) {
                const int col = it / N, e = it - col * N;
                const int b = e % M, rest = e / M;  // rest = q + r u
                const int q = rest % r, u = rest / r;
                if (q > 0 && u > 0) inw[col * N + e] = cmul(inw[col * N + e], twid<INV>(tw, q * u * M));
            }
            __syncthreads();
        }
        const int step = N / r;
        for (int it = tid; it < ncol * N; it += kFFTThreads) {
            const int col = it / N, e = it - col * N;  // e = b + M(u + l s)
            const int b = e % M, rest = e / M;
            const int u = rest % l, s = rest / l;
            const float2 *src = in + col * N + b + M * r * u;
            float2 acc = src[0];
            int idx = 0;
            const int inc = s * step;
            for (int q = 1; q < r; ++q) {
                idx += inc;
                if (idx >= N) idx -= N;
                const float2 v = src[M * q];
                const float2 t = twid<INV>(tw, idx);
                acc.x += v.x * t.x - v.y * t.y;
                acc.y += v.x * t.y + v.y * t.x;
            }
            out[col * N + e] = acc;
        }
    }
}

// Radix-r pass (r odd, >= 11) as two small GEMMs on MFMA.  With the twiddled inputs
// x_q of one butterfly, a_t = x_t + x_{r-t}, b_t = x_t - x_{r-t} (t = 1..h, h = (r-1)/2):
//   T1[s] = sum_t cos(2 pi s t / r) a_t   (s = 0..h),   T2[s] = sum_t sin(2 pi s t / r) b_t
//   y_0 = x_0 + T1[0],  y_s = x_0 + T1[s] -+ i T2[s],  y_{r-s} = x_0 + T1[s] +- i T2[s]
// (upper signs forward).  GEMM rows = s, columns = (butterfly, re/im), K = t; the cos/sin
// operand comes from the ring's twiddle table in LDS at index (s t mod r) N/r.
template <bool INV>
__device__ void fft_pass_mfma(float2 *in, float2 *out, const float2 *tw, int N, int r, int l, int ncol,
                              int tid) {
    const int M = N / (l * r);
    const int h = (r - 1) >> 1;
    const int nbf = N / r;          // butterflies per column
    const int nseq = ncol * nbf;
    // twiddles (only for u > 0) and the pair sums / differences, in place
    for (int it = tid; it < nseq * h; it += kFFTThreads) {
        const int sq = it / h, t = it - sq * h + 1;
        const int col = sq / nbf, bt = sq - col * nbf;
        const int b = bt % M, u = bt / M;
        float2 *x = in + col * N + b + M * r * u;
        float2 xt = x[M * t], xr = x[M * (r - t)];
        if (u > 0) {
            xt = cmul(xt, twid<INV>(tw, t * u * M));
            xr = cmul(xr, twid<INV>(tw, (r - t) * u * M));
        }
        x[M * t] = cadd(xt, xr);
        x[M * (r - t)] = csub(xt, xr);
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, row = lane & 15, grp = lane >> 4;
    const int Mt = (h + 16) >> 4;             // rows 0..h
    const int Nt = (2 * nseq + 15) >> 4;      // real columns
    const int Kt = (h + 3) >> 2;
    const int step = N / r;
    for (int tile = wave; tile < Mt * Nt; tile += kFFTThreads / 64) {
        const int mt = tile % Mt, nt = tile / Mt;
        const int sA = mt * 16 + row;                 // A-operand row
        const int c = nt * 16 + row;                  // B-operand / output column
        const int sq = c >> 1, reim = c & 1;
        const bool cval = sq < nseq;
        int base = 0;
        if (cval) {
            const int col = sq / nbf, bt = sq - col * nbf;
            base = col * N + bt % M + M * r * (bt / M);
        }
        const float *xf = reinterpret_cast<const float *>(in + base) + reim;
        f4 t1 = {0.f, 0.f, 0.f, 0.f}, t2 = {0.f, 0.f, 0.f, 0.f};
        int idx = (sA * (grp + 1)) % r;
        const int inc = (4 * sA) % r;
        for (int ks = 0; ks < Kt; ++ks) {
            const int t = ks * 4 + grp + 1;
            float cs = 0.f, sn = 0.f, av = 0.f, bv = 0.f;
            if (t <= h) {
                if (sA <= h) {
                    const float2 w = tw[idx * step];
                    cs = w.x;
                    sn = -w.y;
                }
                if (cval) {
                    av = xf[2 * M * t];
                    bv = xf[2 * M * (r - t)];
                }
            }
            t1 = mfma_f32(cs, av, t1);
            t2 = mfma_f32(sn, bv, t2);
            idx += inc;
            if (idx >= r) idx -= r;
        }
        // lane holds rows s = 16 mt + 4 grp + i, column c; pair re/im through lane^1
        float2 x0 = make_float2(0.f, 0.f);
        int ob = 0;
        if (cval) {
            x0 = in[base];
            const int col = sq / nbf, bt = sq - col * nbf;
            ob = col * N + bt % M + M * (bt / M);  // out[b + M(u + l s)] = ob + M l s
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float a1 = t1[i], a2 = t2[i];
            const float p1 = __shfl_xor(a1, 1), p2 = __shfl_xor(a2, 1);
            const int s = mt * 16 + 4 * grp + i;
            if (!cval || s > h) continue;
            // re lane (reim 0): T1re = a1, T1im = p1, T2re = a2, T2im = p2; im lane swapped
            const float T1re = reim ? p1 : a1, T1im = reim ? a1 : p1;
            const float T2re = reim ? p2 : a2, T2im = reim ? a2 : p2;
            const float sg = INV ? -1.f : 1.f;
            if (s == 0) {
                if (reim == 0) out[ob] = make_float2(x0.x + T1re, x0.y + T1im);
            } else if (reim == 0) {
                out[ob + M * l * s] = make_float2(x0.x + T1re + sg * T2im, x0.y + T1im - sg * T2re);
            } else {
                out[ob + M * l * (r - s)] = make_float2(x0.x + T1re - sg * T2im, x0.y + T1im + sg * T2re);
            }
        }
    }
}

// Runs all passes; returns the buffer holding the result.
template <bool INV>
__device__ float2 *fft_run(float2 *b0, float2 *b1, const float2 *tw, const int *fac, int nf, int N,
                           int ncol, int tid) {
    int l = 1;
    float2 *in = b0, *out = b1;
    for (int f = 0; f < nf; ++f) {
        const int r = fac[f];
        if (r >= 11)
            fft_pass_mfma<INV>(in, out, tw, N, r, l, ncol, tid);
        else
            fft_pass<INV>(in, out, tw, N, r, l, ncol, tid);
        __syncthreads();
        l *= r;
        float2 *t = in;
        in = out;
        out = t;
    }
    return in;
}

// spectral -> grid, Fourier stage (fourier.jl:215-250): per ring, brfft of g_N and g_S
// as one complex inverse FFT of Z = g_N + i g_S (Hermitian-extended; Im of the DC and
// Nyquist bins ignored as FFTW's c2r does; orders above mt are zero).
__global__ __launch_bounds__(kFFTThreads) void fft_synthesis_kernel(
    DevPlan P, const float4 *__restrict__ fbuf, float *__restrict__ field, long long ld_grid, int K) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int j = P.ring_order[blockIdx.x];
    const int N = P.nlon[j], mt = P.mt[j];
    const int k0 = blockIdx.y * P.kg;
    const int ncol = min(P.kg, K - k0);
    float2 *tw = smem;
    float2 *b0 = smem + N;
    float2 *b1 = b0 + P.kg * N;
    const float2 *gtw = reinterpret_cast<const float2 *>(P.twiddle) + P.tw_off[j];
    for (int i = tid; i < N; i += kFFTThreads) tw[i] = gtw[i];
    const long long fb0 = P.fboff[j];
    const int half = N / 2;
    for (int it = tid; it < ncol * (half + 1); it += kFFTThreads) {
        const int mm = it / ncol, col = it - mm * ncol;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (mm <= mt) v = fbuf[(fb0 + mm) * (long long)K + k0 + col];
        float2 *z = b0 + col * N;
        if (mm == 0) {
            z[0] = make_float2(v.x, v.z);
        } else if (mm < half) {
            z[mm] = make_float2(v.x - v.w, v.y + v.z);
            z[N - mm] = make_float2(v.x + v.w, v.z - v.y);
        } else {
            z[half] = make_float2(v.x, v.z);
        }
    }
    __syncthreads();
    const float2 *res = fft_run<true>(b0, b1, tw, P.fac + j * kMaxFactors, P.nfac[j], N, ncol, tid);
    const long long rn = P.ring_n[j], rs = P.ring_s[j];
    for (int it = tid; it < ncol * N; it += kFFTThreads) {
        const int col = it / N, i = it - col * N;
        const float2 z = res[col * N + i];
        float *f = field + (long long)(k0 + col) * ld_grid;
        f[rn + i] = z.x;
        f[rs + i] = z.y;
    }
}

// grid -> spectral, Fourier stage (fourier.jl:129-166): per ring, rfft of the northern
// and southern ring as one complex FFT of z = f_N + i f_S, separated by symmetry into
// f_N, f_S and stored as (f_N + f_S, f_N - f_S) for orders 0..mt.
__global__ __launch_bounds__(kFFTThreads) void fft_analysis_kernel(
    DevPlan P, const float *__restrict__ field, long long ld_grid, float4 *__restrict__ fbuf, int K) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const int j = P.ring_order[blockIdx.x];
    const int N = P.nlon[j], mt = P.mt[j];
    const int k0 = blockIdx.y * P.kg;
    const int ncol = min(P.kg, K - k0);
    float2 *tw = smem;
    float2 *b0 = smem + N;
    float2 *b1 = b0 + P.kg * N;
    const float2 *gtw = reinterpret_cast<const float2 *>(P.twiddle) + P.tw_off[j];
    for (int i = tid; i < N; i += kFFTThreads) tw[i] = gtw[i];
    const long long rn = P.ring_n[j], rs = P.ring_s[j];
    for (int it = tid; it < ncol * N; it += kFFTThreads) {
        const int col = it / N, i = it - col * N;
        const float *f = field + (long long)(k0 + col) * ld_grid;
        b0[col * N + i] = make_float2(f[rn + i], f[rs + i]);
    }
    __syncthreads();
    const float2 *res = fft_run<false>(b0, b1, tw, P.fac + j * kMaxFactors, P.nfac[j], N, ncol, tid);
    const long long fb0 = P.fboff[j];
    for (int it = tid; it < ncol * (mt + 1); it += kFFTThreads) {
        const int mm = it / ncol, col = it - mm * ncol;
        const float2 p = res[col * N + mm];
        const float2 q = res[col * N + (mm == 0 ? 0 : N - mm)];
        // f_N = (Z_m + conj Z_{N-m})/2, f_S = (Z_m - conj Z_{N-m})/(2i)
        const float nre = 0.5f * (p.x + q.x), nim = 0.5f * (p.y - q.y);
        const float sre = 0.5f * (p.y + q.y), sim = 0.5f * (q.x - p.x);
        fbuf[(fb0 + mm) * (long long)K + k0 + col] = make_float4(nre + sre, nim + sim, nre - sre, nim - sim);
    }
}

// ============================================================ UV from vor, div
// spectral_gradients.jl:331-366: per packed (l, m) a 3-point stencil in l.
__global__ void uv_from_vordiv_kernel(DevPlan P, const float2 *__restrict__ vor,
                                      const float2 *__restrict__ dv, float2 *__restrict__ U,
                                      float2 *__restrict__ V, long long ld, int K, float radius) {
    const long long total = (long long)P.nharm * K;
    for (long long it = blockIdx.x * (long long)blockDim.x + threadIdx.x; it < total;
         it += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(it / P.nharm);
        const int lm = (int)(it - (long long)k * P.nharm);
        const int l = P.l_of[lm];
        const long long base = (long long)k * ld;
        const float x = P.uv_x[lm], u1 = P.uv_1[lm], u2 = P.uv_2[lm];
        const float2 z = vor[base + lm], d = dv[base + lm];
        float2 du = make_float2(0.f, 0.f), dvv = make_float2(0.f, 0.f);
        // i x D and i x ζ
        float2 uu, vv;
        if (lm == 0) {
            float2 zp = make_float2(0.f, 0.f), dp = make_float2(0.f, 0.f);
            if (P.lmax > 1) {
                zp = vor[base + 1];
                dp = dv[base + 1];
            }
            uu = make_float2(-x * d.y + u2 * zp.x, x * d.x + u2 * zp.y);
            vv = make_float2(-x * z.y - u2 * dp.x, x * z.x - u2 * dp.y);
        } else if (l == P.lmax - 1) {
            const float2 zm = vor[base + lm - 1], dm = dv[base + lm - 1];
            uu = make_float2(-u1 * zm.x, -u1 * zm.y);
            vv = make_float2(u1 * dm.x, u1 * dm.y);
        } else {
            const float2 zm = vor[base + lm - 1], dm = dv[base + lm - 1];
            if (l < P.lmax - 2) {
                du = vor[base + lm + 1];
                dvv = dv[base + lm + 1];
            }
            // u2 is 0 on the second-to-last row, so du, dvv only enter below it
            uu = make_float2(-x * d.y + u2 * du.x - u1 * zm.x, x * d.x + u2 * du.y - u1 * zm.y);
            vv = make_float2(-x * z.y + u1 * dm.x - u2 * dvv.x, x * z.x + u1 * dm.y - u2 * dvv.y);
        }
        U[base + lm] = make_float2(uu.x * radius, uu.y * radius);
        V[base + lm] = make_float2(vv.x * radius, vv.y * radius);
    }
}

// ================================================================== launchers

static inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

int launch_legendre_synthesis(const DevPlan &P, int mode, const LegTile *tiles, int ntiles,
                              const void *coeffs, long long ld_spec, void *fbuf, int K,
                              int unscale, void *stream) {
    dim3 grid(ntiles, (K + kColsComplex - 1) / kColsComplex);
    const float2 *c = static_cast<const float2 *>(coeffs);
    float4 *f = static_cast<float4 *>(fbuf);
    if (mode == kFP32)
        legendre_synthesis_kernel<kFP32><<<grid, 256, 0, S(stream)>>>(P, tiles, c, ld_spec, f, K, unscale);
    else if (mode == kFP16x3)
        legendre_synthesis_kernel<kFP16x3><<<grid, 256, 0, S(stream)>>>(P, tiles, c, ld_spec, f, K, unscale);
    else
        legendre_synthesis_kernel<kBF16x3><<<grid, 256, 0, S(stream)>>>(P, tiles, c, ld_spec, f, K, unscale);
    return (int)hipGetLastError();
}

int launch_legendre_analysis(const DevPlan &P, int mode, const LegTile *tiles, int ntiles,
                             const void *fbuf, void *coeffs, long long ld_spec, int K,
                             void *stream) {
    dim3 grid(ntiles, (K + kColsComplex - 1) / kColsComplex);
    const float4 *f = static_cast<const float4 *>(fbuf);
    float2 *c = static_cast<float2 *>(coeffs);
    if (mode == kFP32)
        legendre_analysis_kernel<kFP32><<<grid, 256, 0, S(stream)>>>(P, tiles, f, c, ld_spec, K);
    else if (mode == kFP16x3)
        legendre_analysis_kernel<kFP16x3><<<grid, 256, 0, S(stream)>>>(P, tiles, f, c, ld_spec, K);
    else
        legendre_analysis_kernel<kBF16x3><<<grid, 256, 0, S(stream)>>>(P, tiles, f, c, ld_spec, K);
    return (int)hipGetLastError();
}

int launch_fft_synthesis(const DevPlan &P, const void *fbuf, void *field, long long ld_grid,
                         int K, void *stream) {
    if (P.fft_smem > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fft_synthesis_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, P.fft_smem);
    dim3 grid(P.n, (K + P.kg - 1) / P.kg);
    fft_synthesis_kernel<<<grid, kFFTThreads, P.fft_smem, S(stream)>>>(
        P, static_cast<const float4 *>(fbuf), static_cast<float *>(field), ld_grid, K);
    return (int)hipGetLastError();
}

int launch_fft_analysis(const DevPlan &P, const void *field, long long ld_grid, void *fbuf,
                        int K, void *stream) {
    if (P.fft_smem > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fft_analysis_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, P.fft_smem);
    dim3 grid(P.n, (K + P.kg - 1) / P.kg);
    fft_analysis_kernel<<<grid, kFFTThreads, P.fft_smem, S(stream)>>>(
        P, static_cast<const float *>(field), ld_grid, static_cast<float4 *>(fbuf), K);
    return (int)hipGetLastError();
}

int launch_uv_from_vordiv(const DevPlan &P, const void *vor, const void *div, void *U,
                          void *V, long long ld, int K, float radius, void *stream) {
    const long long total = (long long)P.nharm * K;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    uv_from_vordiv_kernel<<<blocks, 256, 0, S(stream)>>>(
        P, static_cast<const float2 *>(vor), static_cast<const float2 *>(div),
        static_cast<float2 *>(U), static_cast<float2 *>(V), ld, K, radius);
    return (int)hipGetLastError();
}

}  // namespace sht
