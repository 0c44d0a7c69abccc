// gfx950 ring FFT of the spherical harmonic transform (Fourier stage).
//
// One complex FFT of length nlon per ring and column carries the northern ring in the
// real part and the mirrored southern ring in the imaginary part: the per-ring
// rfft / brfft pairs of SpeedyTransforms/src/fourier.jl:27-77,129-250 for both hemispheres
// at once.  Mixed-radix Stockham autosort in LDS, decimation in time.  With
// N = r_1 r_2 ... r_s, pass p of radix r (l = r_1...r_{p-1}, M = N/(l r)) maps, for
// b in [0,M), u in [0,l):
//   in[b + M(q + r u)], q < r   ->   out[b + M(u + l s)], s < r
//   out = sum_q (w_N^{q u M} in_q) w_r^{q s},   w_N = exp(-/+ 2 pi i / N).
// Every pass runs in place in one LDS buffer.  A first-pass prime >= 17 (the octahedral
// rings nlon = 16 + 4j have prime factors up to 193 at T255) runs across the workgroup as
// two small GEMMs on MFMA; all other passes (radices 2..13, register butterflies) run
// across the workgroup with one thread per butterfly and group of columns.
#include "sht_device.h"

namespace sht {

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

__device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ inline float2 cfma(float2 a, float s, float2 c) { return make_float2(fmaf(a.x, s, c.x), fmaf(a.y, s, c.y)); }

// DFT of R values in place (R = 2, 3, 4, 5, 7)
template <int R, bool INV>
__device__ inline void bfly(float2 *a) {
    if constexpr (R == 2) {
        const float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if constexpr (R == 4) {
        const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const float2 s13 = cadd(a[1], a[3]), d13 = mul_mi<INV>(csub(a[1], a[3]));
        a[0] = cadd(s02, s13);
        a[1] = cadd(d02, d13);
        a[2] = csub(s02, s13);
        a[3] = csub(d02, d13);
    } else if constexpr (R == 3) {
        const float s3 = 0.86602540378443864676f;
        const float2 s = cadd(a[1], a[2]);
        const float2 d = mul_mi<INV>(csub(a[1], a[2]));
        const float2 base = cfma(s, -0.5f, a[0]);
        a[0] = cadd(a[0], s);
        a[1] = cfma(d, s3, base);
        a[2] = cfma(d, -s3, base);
    } else if constexpr (R == 5) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 s14 = cadd(a[1], a[4]), d14 = mul_mi<INV>(csub(a[1], a[4]));
        const float2 s23 = cadd(a[2], a[3]), d23 = mul_mi<INV>(csub(a[2], a[3]));
        const float2 a0 = a[0];
        const float2 b1 = cfma(s23, c2, cfma(s14, c1, a0));
        const float2 b2 = cfma(s23, c1, cfma(s14, c2, a0));
        const float2 e1 = cfma(d23, s2, cscale(d14, s1));
        const float2 e2 = cfma(d23, -s1, cscale(d14, s2));
        a[0] = cadd(a0, cadd(s14, s23));
        a[1] = cadd(b1, e1);
        a[4] = csub(b1, e1);
        a[2] = cadd(b2, e2);
        a[3] = csub(b2, e2);
    } else {  // R == 7
        const float c1 = 0.62348980185873353053f, c2 = -0.22252093395631440429f, c3 = -0.90096886790241912624f;
        const float s1 = 0.78183148246802980871f, s2 = 0.97492791218182360702f, s3 = 0.43388373911755812048f;
        const float2 a0 = a[0];
        const float2 p1 = cadd(a[1], a[6]), q1 = mul_mi<INV>(csub(a[1], a[6]));
        const float2 p2 = cadd(a[2], a[5]), q2 = mul_mi<INV>(csub(a[2], a[5]));
        const float2 p3 = cadd(a[3], a[4]), q3 = mul_mi<INV>(csub(a[3], a[4]));
        const float2 b1 = cfma(p3, c3, cfma(p2, c2, cfma(p1, c1, a0)));
        const float2 b2 = cfma(p3, c1, cfma(p2, c3, cfma(p1, c2, a0)));
        const float2 b3 = cfma(p3, c2, cfma(p2, c1, cfma(p1, c3, a0)));
        const float2 e1 = cfma(q3, s3, cfma(q2, s2, cscale(q1, s1)));
        const float2 e2 = cfma(q3, -s1, cfma(q2, -s3, cscale(q1, s2)));
        const float2 e3 = cfma(q3, s2, cfma(q2, -s1, cscale(q1, s3)));
        a[0] = cadd(a0, cadd(p1, cadd(p2, p3)));
        a[1] = cadd(b1, e1);
        a[6] = csub(b1, e1);
        a[2] = cadd(b2, e2);
        a[5] = csub(b2, e2);
        a[3] = cadd(b3, e3);
        a[4] = csub(b3, e3);
    }
}

// DFT of an odd prime number R of values in place (R = 11, 13), by the symmetric
// pairs a_t +- a_{R-t}; cos / sin of 2 pi k / R come from the ring's table tw at k N/R.
template <int R, bool INV>
__device__ inline void bfly_odd(float2 *a, const float2 *tw, int stepR) {
    constexpr int H = (R - 1) / 2;
    float cs[H + 1], sn[H + 1];
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        const float2 w = tw[k * stepR];
        cs[k] = w.x;
        sn[k] = -w.y;
    }
    float2 p[H], q[H];
    float2 y0 = a[0];
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        p[t - 1] = cadd(a[t], a[R - t]);
        q[t - 1] = mul_mi<INV>(csub(a[t], a[R - t]));
        y0 = cadd(y0, p[t - 1]);
    }
#pragma unroll
    for (int s = 1; s <= H; ++s) {
        float2 b = a[0], e = make_float2(0.f, 0.f);
#pragma unroll
        for (int t = 1; t <= H; ++t) {
            const int k = (s * t) % R;
            const float c = k <= H ? cs[k] : cs[R - k];
            const float sg = k <= H ? sn[k] : -sn[R - k];
            b = cfma(p[t - 1], c, b);
            e = cfma(q[t - 1], sg, e);
        }
        a[s] = cadd(b, e);
        a[R - s] = csub(b, e);
    }
    a[0] = y0;
}

// multiply by w8^k = exp(-+ i pi k / 4), k = 1, 2, 3
template <int K, bool INV>
__device__ inline float2 mul_w8(float2 a) {
    constexpr float h = 0.70710678118654752440f;
    if constexpr (K == 2) return mul_mi<INV>(a);
    // forward: w8 = (1 - i) h, w8^3 = (-1 - i) h; inverse: conjugates
    const float2 m = mul_mi<INV>(a);  // a * w8^2
    if constexpr (K == 1) return make_float2(h * (a.x + m.x), h * (a.y + m.y));
    return make_float2(h * (m.x - a.x), h * (m.y - a.y));
}

// radix 8: two radix-4 halves (even / odd inputs) joined by w8^k
template <bool INV>
__device__ inline void bfly8(float2 *a) {
    float2 e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
    bfly<4, INV>(e);
    bfly<4, INV>(o);
    o[1] = mul_w8<1, INV>(o[1]);
    o[2] = mul_w8<2, INV>(o[2]);
    o[3] = mul_w8<3, INV>(o[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = cadd(e[k], o[k]);
        a[k + 4] = csub(e[k], o[k]);
    }
}

// radix 16: four radix-4 DFTs over q = t mod 4, twiddles w16^{q k1}, radix-4 across q
template <bool INV>
__device__ inline void bfly16(float2 *a) {
    // cos / sin of 2 pi k / 16, k = 0..9 (only q k1 <= 9 occurs)
    const float C[10] = {1.f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f, 0.f,
                         -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f, -1.f,
                         -0.92387953251128675613f};
    const float S[10] = {0.f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f, 1.f,
                         0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f, 0.f,
                         -0.38268343236508977173f};
    float2 b[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        b[q][0] = a[q];
        b[q][1] = a[q + 4];
        b[q][2] = a[q + 8];
        b[q][3] = a[q + 12];
        bfly<4, INV>(b[q]);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = q * k1;
            const float2 v = b[q][k1];
            if (e == 0) {
                c[q] = v;
            } else {
                const float sn = INV ? S[e] : -S[e];  // w16^e = cos - i sin (forward)
                c[q] = make_float2(v.x * C[e] - v.y * sn, v.x * sn + v.y * C[e]);
            }
        }
        bfly<4, INV>(c);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = c[k2];
    }
}

template <int R, bool INV>
__device__ inline void bfly_any(float2 *a, const float2 *tw, int stepR) {
    if constexpr (R == 8)
        bfly8<INV>(a);
    else if constexpr (R == 16)
        bfly16<INV>(a);
    else if constexpr (R <= 7)
        bfly<R, INV>(a);
    else
        bfly_odd<R, INV>(a, tw, stepR);
}

// A pass of radix R over all columns of N <= kFFTMaxRing points, run by the workgroup.
// Work item = (column group of CB columns, butterfly t); items go to threads in order, so
// neighbouring lanes take neighbouring butterflies of one column group.  The index
// arithmetic and the twiddles of a butterfly serve all CB columns of its item.  Every
// thread reads all its items into registers, the workgroup synchronises, then the results
// are written in place.  Columns past the call's last column hold scratch values that are
// transformed along and never stored.
// Radices 11 and 13 (one column per item) run as two sub-passes of two columns each, so
// that a thread holds at most one butterfly; radix 16 (N / 16 butterflies per column, at
// most 256 items over four columns) needs no split.
template <int R>
struct PassCols {
    static constexpr int CB = R <= 7 ? 4 : R <= 8 ? 2 : 1;  // columns per work item
    static constexpr int SUB = R <= 8 || R == 16 ? 1 : 2;   // column sub-passes
    static constexpr int NCG = kFFTCols / CB / SUB;         // column groups per sub-pass
    static constexpr int MAXI = (kFFTMaxRing / R * NCG + kFFTThreads - 1) / kFFTThreads;
};

template <int R, bool INV>
__device__ void fft_subpass_wg(float2 *buf, const float2 *tw, int N, const FFTPass &pd, int tid) {
    constexpr int CB = PassCols<R>::CB, NCG = PassCols<R>::NCG, MAXI = PassCols<R>::MAXI;
    const int M = pd.M, l = pd.l;
    const int nb = N / R;  // butterflies per column
    const int items = nb * NCG;
    const FastDiv divM(M, pd.mulM);
    // item -> (column group, butterfly), NCG <= 4
    auto split = [&](int it, int &cg, int &t) {
        cg = 0;
#pragma unroll
        for (int k = 1; k < NCG; ++k) cg += it >= k * nb ? 1 : 0;
        t = it - cg * nb;
    };
    float2 a[MAXI][CB][R];
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
        const int it = tid + kFFTThreads * i;
        if (it < items) {
            int cg, t;
            split(it, cg, t);
            const int u = divM.div(t), b = t - u * M;
            const float2 *src = buf + cg * CB * N + b + M * R * u;
#pragma unroll
            for (int cc = 0; cc < CB; ++cc)
#pragma unroll
                for (int q = 0; q < R; ++q) a[i][cc][q] = src[cc * N + M * q];
        }
    }
    __syncthreads();
    const int ls = M * l;  // output stride between s
    const int stepR = N / R;
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
        const int it = tid + kFFTThreads * i;
        if (it >= items) continue;
        int cg, t;
        split(it, cg, t);
        const int u = divM.div(t), b = t - u * M;
        if (u > 0) {
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const float2 w = twid<INV>(tw, q * u * M);
#pragma unroll
                for (int cc = 0; cc < CB; ++cc) a[i][cc][q] = cmul(a[i][cc][q], w);
            }
        }
        float2 *dst = buf + cg * CB * N + b + M * u;
#pragma unroll
        for (int cc = 0; cc < CB; ++cc) {
            bfly_any<R, INV>(a[i][cc], tw, stepR);
#pragma unroll
            for (int s = 0; s < R; ++s) dst[cc * N + ls * s] = a[i][cc][s];
        }
    }
    __syncthreads();
}

template <int R, bool INV>
__device__ void fft_pass_wg(float2 *buf, const float2 *tw, int N, const FFTPass &pd, int tid) {
#pragma unroll
    for (int sp = 0; sp < PassCols<R>::SUB; ++sp)
        fft_subpass_wg<R, INV>(buf + sp * (kFFTCols / PassCols<R>::SUB) * N, tw, N, pd, tid);
}

// First pass (l = 1) of an odd prime radix r >= 11 as two small GEMMs on MFMA, in place.
// With the inputs x_q = buf[b + M q] of butterfly b, a_t = x_t + x_{r-t},
// b_t = x_t - x_{r-t} (t = 1..h, h = (r-1)/2):
//   T1[s] = sum_t cos(2 pi s t / r) a_t   (s = 0..h),   T2[s] = sum_t sin(2 pi s t / r) b_t
//   y_0 = x_0 + T1[0],  y_s = x_0 + T1[s] -+ i T2[s],  y_{r-s} = x_0 + T1[s] +- i T2[s]
// (upper signs forward), y_s -> buf[b + M s].  GEMM rows = s (row tiles of 16), columns =
// (butterfly, re/im) (column tiles of 16), K = t in chunks of 32.
//
// Products in three-part BF16 (v_mfma_f32_16x16x32_bf16, 16 cycles per 32-wide K chunk,
// against 8 FP32 MFMAs 16x16x4 of 32 cycles): cos / sin come from the plan's table for r
// (hi + mid + lo of the FP64 values, sht_plan.cpp: prime_tables); a_t and b_t are split the
// same way once per chunk and wave (x = hi + mid + lo exactly: BF16 keeps FP32's exponent
// range, so no scaling); the six products whose parts sum to at most 2 accumulate in FP32,
// smallest first, so the pass keeps FP32-level accuracy in every precision mode.
//
// Column tiles go to the waves in rounds of four; with fewer column tiles left, the waves of
// one column tile split its row tiles (at most kPrimeRowTiles per wave: rings of N = 4p
// points have M = N / r >= 4, see sht_plan.cpp: factorize).  All row tiles of a round's
// column tiles are computed before the workgroup synchronises and any output is written
// (a butterfly's outputs overwrite its inputs, which the other row tiles still read).
// tmax: the last pair t with nonzero inputs (a synthesis reads only the orders m <= mt and
// their mirrors N - m: x_t, x_{r-t} vanish for t > (mt + M - 1) / M); h otherwise.
__device__ inline void split3_bf16(const float (&x)[8], b8 &hi, b8 &mid, b8 &lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)x[e];
        const float r1 = x[e] - (float)h;  // exact
        const __bf16 m = (__bf16)r1;
        hi[e] = h;
        mid[e] = m;
        lo[e] = (__bf16)(r1 - (float)m);
    }
}

// c += a b over the parts (i, j) with i + j <= 2, smallest first
__device__ inline f4 mma_bf16x3(const b8 &ah, const b8 &am, const b8 &al, const b8 &bh, const b8 &bm, const b8 &bl,
                                f4 c) {
    c = Split<kBF16x3>::mma(al, bh, c);
    c = Split<kBF16x3>::mma(ah, bl, c);
    c = Split<kBF16x3>::mma(am, bm, c);
    c = Split<kBF16x3>::mma(am, bh, c);
    c = Split<kBF16x3>::mma(ah, bm, c);
    return Split<kBF16x3>::mma(ah, bh, c);
}

template <bool INV>
__device__ void fft_pass_mfma_first(float2 *buf, const SHT_GLOBAL b8 *tab, int N, const FFTPass &pd, int ncol,
                                    int tmax, int tid) {
    const int r = pd.R, M = pd.M;
    const int h = (r - 1) >> 1;
    const int nseq = ncol * M;  // butterflies over all columns
    constexpr int NW = kFFTThreads / 64;
    // the wave index is uniform per wave: keep it (and all tile ownership) scalar
    const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int row = lane & 15, grp = lane >> 4, ri = row & 1;
    const int Mt = (h + 16) >> 4;         // row tiles (rows 0..h)
    const int Nt = (2 * nseq + 15) >> 4;  // column tiles over (butterfly, re/im)
    const int KC = (h + 31) >> 5;         // K chunks of the table
    const int KU = (tmax + 31) >> 5;      // K chunks with nonzero inputs
    const FastDiv divM(M, pd.mulM);
    const float *bf = reinterpret_cast<const float *>(buf);
    for (int n0 = 0; n0 < Nt; n0 += NW) {  // uniform: every wave takes part in every round's barriers
        const int nc = min(NW, Nt - n0);   // column tiles this round
        const int G = NW / nc;             // waves per column tile
        const int nt = n0 + wave % nc, r0 = wave / nc;
        const int nrt = wave < nc * G ? min((Mt - r0 + G - 1) / G, kPrimeRowTiles) : 0;  // uniform per wave
        f4 t1[kPrimeRowTiles], t2[kPrimeRowTiles];
#pragma unroll
        for (int i = 0; i < kPrimeRowTiles; ++i) t1[i] = t2[i] = f4{0.f, 0.f, 0.f, 0.f};
        int ob = -1;
        float2 z0 = make_float2(0.f, 0.f);
        if (nrt > 0) {
            const int c = nt * 16 + row, sq = c >> 1;
            int base = 0;  // columns past the data read column 0; their outputs are dropped
            if (sq < nseq) {
                const int col = divM.div(sq);
                base = col * N + (sq - col * M);
                ob = base;
                z0 = buf[base];  // x_0, read before any output of the round is written
            }
            for (int kc = 0; kc < KU; ++kc) {
                float xs[8], xd[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int t = 32 * kc + 8 * grp + e + 1;
                    const bool ok = t <= tmax;
                    const int tt = ok ? t : tmax;
                    const float xa = bf[2 * (base + M * tt) + ri], xb = bf[2 * (base + M * (r - tt)) + ri];
                    xs[e] = ok ? xa + xb : 0.f;
                    xd[e] = ok ? xa - xb : 0.f;
                }
                b8 sh, sm, sl, dh, dm, dl;
                split3_bf16(xs, sh, sm, sl);
                split3_bf16(xd, dh, dm, dl);
#pragma unroll
                for (int i = 0; i < kPrimeRowTiles; ++i) {
                    if (i >= nrt) break;  // uniform per wave
                    const SHT_GLOBAL b8 *a = tab + ((long long)((r0 + i * G) * KC + kc) * 64 + lane) * 6;
                    const b8 ch = a[0], cm = a[1], cl = a[2];
                    t1[i] = mma_bf16x3(ch, cm, cl, sh, sm, sl, t1[i]);
                    const b8 nh = a[3], nm = a[4], nl = a[5];
                    t2[i] = mma_bf16x3(nh, nm, nl, dh, dm, dl, t2[i]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kPrimeRowTiles; ++i) {
            if (i >= nrt) break;  // uniform per wave
            const int mt = r0 + i * G;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                // the partner lane (c ^ 1) holds the other part of the same butterfly
                const float a1 = t1[i][e], a2 = t2[i][e];
                const float p1 = lane_xor1(a1), p2 = lane_xor1(a2);
                const int s = mt * 16 + 4 * grp + e;
                if (ob < 0 || s > h) continue;
                const float T1re = ri ? p1 : a1, T1im = ri ? a1 : p1;
                const float T2re = ri ? p2 : a2, T2im = ri ? a2 : p2;
                const float sg = INV ? -1.f : 1.f;
                if (s == 0) {
                    if (ri == 0) buf[ob] = make_float2(z0.x + T1re, z0.y + T1im);
                } else if (ri == 0) {
                    buf[ob + M * s] = make_float2(z0.x + T1re + sg * T2im, z0.y + T1im - sg * T2re);
                } else {
                    buf[ob + M * (r - s)] = make_float2(z0.x + T1re - sg * T2im, z0.y + T1im + sg * T2re);
                }
            }
        }
        __syncthreads();
    }
}

// A pass descriptor read field by field (structures do not copy through global pointers).
__device__ __forceinline__ FFTPass load_pass(const SHT_GLOBAL FFTPass *p) {
    const SHT_GLOBAL int *q = reinterpret_cast<const SHT_GLOBAL int *>(p);
    FFTPass d;
    d.R = q[0];
    d.l = q[1];
    d.M = q[2];
    d.mulM = (unsigned)q[3];
    d.mulR = (unsigned)q[4];
    return d;
}

// Runs all passes in place on buf (kFFTCols columns of N points, the first ncol of them
// data): a first-pass MFMA radix, then the register passes, all across the workgroup.
template <bool INV>
__device__ __attribute__((always_inline)) void fft_run(const DevPlan &P, const FFTRing &R, float2 *buf,
                                                        const float2 *tw, int N, int mt, int ncol, int tid) {
    const SHT_GLOBAL FFTPass *passes = as_global(P.passes) + R.passes;
    const int nf = R.nf;
    int f = 0;
    if (R.prime) {  // the MFMA first pass (its descriptor and table offset come with the ring)
        FFTPass p0;
        p0.R = R.prime;
        p0.l = 1;
        p0.M = R.M0;
        p0.mulM = R.mulM0;
        p0.mulR = R.mulR0;
        const SHT_GLOBAL b8 *tab = reinterpret_cast<const SHT_GLOBAL b8 *>(as_global(P.prime_tab)) + R.prime_off;
        const int h = (p0.R - 1) >> 1;
        const int tmax = INV ? min(h, (mt + p0.M - 1) / p0.M) : h;
        fft_pass_mfma_first<INV>(buf, tab, N, p0, ncol, max(tmax, 1), tid);
        f = 1;
    }
    for (int g = f; g < nf; ++g) {
        const FFTPass pd = load_pass(passes + g);
        switch (pd.R) {
            case 16: fft_pass_wg<16, INV>(buf, tw, N, pd, tid); break;
            case 8: fft_pass_wg<8, INV>(buf, tw, N, pd, tid); break;
            case 2: fft_pass_wg<2, INV>(buf, tw, N, pd, tid); break;
            case 3: fft_pass_wg<3, INV>(buf, tw, N, pd, tid); break;
            case 4: fft_pass_wg<4, INV>(buf, tw, N, pd, tid); break;
            case 5: fft_pass_wg<5, INV>(buf, tw, N, pd, tid); break;
            case 7: fft_pass_wg<7, INV>(buf, tw, N, pd, tid); break;
            case 11: fft_pass_wg<11, INV>(buf, tw, N, pd, tid); break;
            default: fft_pass_wg<13, INV>(buf, tw, N, pd, tid); break;  // the plan admits no other radix
        }
    }
}

// Twiddles of ring j into LDS with all loads of a thread in flight together.
__device__ inline void load_twiddles(const DevPlan &P, long long tw_off, int N, float2 *tw, int tid) {
    const SHT_GLOBAL float2 *gtw = as_global(reinterpret_cast<const float2 *>(P.twiddle)) + tw_off;
    constexpr int U = (kFFTMaxRing + kFFTThreads - 1) / kFFTThreads;
    float2 v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const int i = q * kFFTThreads + tid;
        if (i < N) v[q] = gload(gtw + i);
    }
    __builtin_amdgcn_sched_barrier(0);  // every load in flight before the first LDS store
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const int i = q * kFFTThreads + tid;
        if (i < N) tw[i] = v[q];
    }
}

// The ring record of a rank, read field by field (structures do not copy through global
// pointers): one level of loads after the rank is known.
__device__ __forceinline__ FFTRing load_ring(const SHT_GLOBAL FFTRing *p) {
    const SHT_GLOBAL int *q = reinterpret_cast<const SHT_GLOBAL int *>(p);
    const SHT_GLOBAL long long *w = reinterpret_cast<const SHT_GLOBAL long long *>(p);
    FFTRing r;
    r.j = q[0];
    r.N = q[1];
    r.mt = q[2];
    r.mte = q[3];
    r.nf = q[4];
    r.prime = q[5];
    r.M0 = q[6];
    r.mulM0 = (unsigned)q[7];
    r.mulR0 = (unsigned)q[8];
    r.tw_off = w[6];
    r.fboff = w[7];
    r.ring_n = w[8];
    r.ring_s = w[9];
    r.prime_off = w[10];
    r.passes = w[11];
    return r;
}

// Per-workgroup launch geometry of the FFT kernels: the ring's record, first column k0,
// columns ncol.
struct FFTTile {
    FFTRing R;
    int k0, ncol;
};

// Single-call launch orders, rings heaviest first.  Paired (synthesis, which reads the
// Fourier rows): units (ring rank r, column-group pair p) go round-robin to the XCDs and
// the two column groups of a unit, which share the 128-byte lines of the Fourier rows, run
// back to back on the same XCD: block b -> XCD b % 8, slot s = b / 8, unit
// (s / 2) * 8 + b % 8, column group 2 p + s % 2.  Otherwise (analysis): block b -> ring
// rank b / kgs, column group b % kgs.  Batch tiles carry the rank (m) and column group.
// rings: the plan's ring records (single calls pass them in place of a tile list).
// Returns false for blocks past the end.
template <bool BATCH>
__device__ inline bool fft_tile(const DevPlan &P, const SHT_GLOBAL FFTRing *rings, const Tile &T, int K, bool paired,
                                FFTTile &t) {
    const int kgs = (K + kFFTCols - 1) / kFFTCols;
    int cg, r;
    if (BATCH) {
        if (T.m < 0) return false;
        r = T.m;
        cg = T.blk;
    } else {
        const int b = blockIdx.x;
        if (paired) {
            // every column group of ring rank r on XCD r % 8 (blocks go round-robin over the
            // XCDs), ranks heaviest first on each: the ring's twiddles and tables are fetched
            // into one L2, and the column groups sharing a line of Fourier rows meet there
            const int s = b / kXCDs;
            r = (s / kgs) * kXCDs + b % kXCDs;
            cg = s - (s / kgs) * kgs;
        } else {
            r = b / kgs;
            cg = b - r * kgs;
        }
        if (r >= P.n || cg >= kgs) return false;
    }
    t.R = load_ring(rings + r);
    t.k0 = cg * kFFTCols;
    t.ncol = min(kFFTCols, K - t.k0);
    return true;
}

// Fourier rows are read and written with thread -> (order mm = tid/4 + 64 u, column tid%4).
constexpr int kFFTRowsPerPass = kFFTThreads / kFFTCols;
constexpr int kFFTRowPasses = (kFFTMaxRing / 2 + kFFTRowsPerPass) / kFFTRowsPerPass;

// spectral -> grid, Fourier stage (fourier.jl:215-250): per ring, brfft of g_N and g_S as
// one complex inverse FFT of Z = g_N + i g_S (Hermitian-extended; Im of the DC and
// Nyquist bins ignored as FFTW's c2r does; orders above mt are zero).  ACC: the pullback
// that adds into the caller's grid (its own instantiation, so the plain transform keeps
// its registers).
template <bool BATCH, bool ACC = false>
__global__ __launch_bounds__(kFFTThreads, 5) void fft_synthesis_kernel(const DevPlan *__restrict__ plans,
                                                                    CallArgs single,
                                                                    const CallArgs *__restrict__ calls,
                                                                    const Tile *__restrict__ tiles) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const Tile T = BATCH ? tiles[blockIdx.x] : Tile{0, 0, 0, 0};
    const CallArgs C = BATCH ? calls[T.call] : single;
    const DevPlan &P = plans[C.plan];
    const int K = C.K;
    const SHT_GLOBAL FFTRing *rings =
        BATCH ? as_global(P.rings) : as_global(reinterpret_cast<const FFTRing *>(tiles));
    FFTTile G;
    if (!fft_tile<BATCH>(P, rings, T, K, C.paired != 0, G)) return;  // padding of the launch order
    const FFTRing &R = G.R;
    const int N = R.N, mt = R.mt, ncol = G.ncol;
    float2 *tw = smem;
    float2 *b0 = smem + N;
    load_twiddles(P, R.tw_off, N, tw, tid);
    {
        const SHT_GLOBAL float4 *fb =
            as_global(static_cast<const float4 *>(C.fbuf)) + R.fboff * (long long)K + G.k0;
        const int half = N / 2;
        const int col = tid % kFFTCols, m0 = tid / kFFTCols;
        float4 v[kFFTRowPasses];
#pragma unroll
        for (int u = 0; u < kFFTRowPasses; ++u) {
            const int mm = m0 + kFFTRowsPerPass * u;
            v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (col < ncol && mm <= mt) v[u] = gload(fb + (unsigned)(mm * K + col));
        }
        __builtin_amdgcn_sched_barrier(0);  // every load in flight before the first LDS store
        if (C.adj) {  // adjoint of the analysis (grid2spec pullback): rows times ΔΩ_j / s_m
            const float wj = as_global(P.domega)[R.j];
#pragma unroll
            for (int u = 0; u < kFFTRowPasses; ++u) {
                const int mm = m0 + kFFTRowsPerPass * u;
                const float sw = mm == 0 || 2 * mm == N ? wj : 0.5f * wj;
                v[u] = make_float4(v[u].x * sw, v[u].y * sw, v[u].z * sw, v[u].w * sw);
            }
        }
        float2 *z = b0 + col * N;
#pragma unroll
        for (int u = 0; u < kFFTRowPasses; ++u) {
            const int mm = m0 + kFFTRowsPerPass * u;
            if (col >= ncol || mm > half) continue;
            const float4 w = v[u];
            if (mm == 0) {
                z[0] = make_float2(w.x, w.z);
            } else if (mm < half) {
                z[mm] = make_float2(w.x - w.w, w.y + w.z);
                z[N - mm] = make_float2(w.x + w.w, w.z - w.y);
            } else {
                z[half] = make_float2(w.x, w.z);
            }
        }
    }
    __syncthreads();
    fft_run<true>(P, R, b0, tw, N, mt, ncol, tid);
    __syncthreads();
    {
        SHT_GLOBAL float *field = as_global(static_cast<float *>(C.out));
        const long long ld = C.ld_out;
        const long long rn = R.ring_n, rs = R.ring_s;
        constexpr int U = (kFFTMaxRing + kFFTThreads - 1) / kFFTThreads;
#pragma unroll
        for (int c = 0; c < kFFTCols; ++c) {
            if (c >= ncol) break;  // uniform
            SHT_GLOBAL float *fn = field + (long long)(G.k0 + c) * ld + rn;
            SHT_GLOBAL float *fs = field + (long long)(G.k0 + c) * ld + rs;
            const float2 *z = b0 + c * N;
            if constexpr (ACC) {
                // pullback: add into the caller's cotangent; the column's old values are all
                // read before its first store, so the reads are in flight together
                float on[U], os[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int i = min(u * kFFTThreads + tid, N - 1);
                    on[u] = fn[(unsigned)i];
                    os[u] = fs[(unsigned)i];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int i = u * kFFTThreads + tid;
                    if (i < N) {
                        const float2 v = z[i];
                        fn[(unsigned)i] = on[u] + v.x;
                        fs[(unsigned)i] = os[u] + v.y;
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int i = u * kFFTThreads + tid;
                    if (i < N) {
                        const float2 v = z[i];
                        fn[(unsigned)i] = v.x;
                        fs[(unsigned)i] = v.y;
                    }
                }
            }
        }
    }
}

// grid -> spectral, Fourier stage (fourier.jl:129-166): per ring, rfft of the northern and
// southern ring as one complex FFT of z = f_N + i f_S, separated by symmetry into f_N, f_S
// and stored as (f_N + f_S, f_N - f_S) for orders 0..mt.
template <bool BATCH>
__global__ __launch_bounds__(kFFTThreads, 5) void fft_analysis_kernel(const DevPlan *__restrict__ plans,
                                                                   CallArgs single,
                                                                   const CallArgs *__restrict__ calls,
                                                                   const Tile *__restrict__ tiles) {
    extern __shared__ float2 smem[];
    const int tid = threadIdx.x;
    const Tile T = BATCH ? tiles[blockIdx.x] : Tile{0, 0, 0, 0};
    const CallArgs C = BATCH ? calls[T.call] : single;
    const DevPlan &P = plans[C.plan];
    const int K = C.K;
    const SHT_GLOBAL FFTRing *rings =
        BATCH ? as_global(P.rings) : as_global(reinterpret_cast<const FFTRing *>(tiles));
    FFTTile G;
    if (!fft_tile<BATCH>(P, rings, T, K, false, G)) return;  // padding of the launch order
    const FFTRing &R = G.R;
    const int N = R.N, mt = R.mt, ncol = G.ncol;
    float2 *tw = smem;
    float2 *b0 = smem + N;
    load_twiddles(P, R.tw_off, N, tw, tid);
    {
        const SHT_GLOBAL float *field = as_global(static_cast<const float *>(C.in));
        const long long ld = C.ld_in;
        const long long rn = R.ring_n, rs = R.ring_s;
        constexpr int U = (kFFTMaxRing + kFFTThreads - 1) / kFFTThreads;
        // all loads of the thread in flight before the first LDS store (columns past ncol
        // re-read column 0 and are not stored)
        float vn[kFFTCols][U], vs[kFFTCols][U];
#pragma unroll
        for (int c = 0; c < kFFTCols; ++c) {
            const int cc = c < ncol ? c : 0;
            const SHT_GLOBAL float *fn = field + (long long)(G.k0 + cc) * ld + rn;
            const SHT_GLOBAL float *fs = field + (long long)(G.k0 + cc) * ld + rs;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned i = min(u * kFFTThreads + tid, N - 1);
                vn[c][u] = fn[i];
                vs[c][u] = fs[i];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < kFFTCols; ++c) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = u * kFFTThreads + tid;
                if (c < ncol && i < N) b0[c * N + i] = make_float2(vn[c][u], vs[c][u]);
            }
        }
    }
    __syncthreads();
    fft_run<false>(P, R, b0, tw, N, mt, ncol, tid);
    __syncthreads();
    {
        SHT_GLOBAL float4 *fb = as_global(static_cast<float4 *>(C.fbuf)) + R.fboff * (long long)K + G.k0;
        const int col = tid % kFFTCols, m0 = tid / kFFTCols;
        const float2 *z = b0 + col * N;
        // adjoint of the synthesis (spec2grid pullback): rows times s_m (1/cos lat) / ΔΩ_j
        const float wj = C.adj ? (C.unscale ? as_global(P.coslat_inv)[R.j] : 1.f) / as_global(P.domega)[R.j] : 1.f;
        // rows mt < m <= mte (orders that rings nearer the pole keep under a non-monotonic
        // shortcut) are written as zeros: the Legendre analysis sums every ring from the
        // first one of an order on
        const int mte = R.mte;
#pragma unroll
        for (int u = 0; u < kFFTRowPasses; ++u) {
            const int mm = m0 + kFFTRowsPerPass * u;
            if (col >= ncol || mm > mte) continue;
            if (mm > mt) {
                gstore(fb + (unsigned)(mm * K + col), make_float4(0.f, 0.f, 0.f, 0.f));
                continue;
            }
            const float2 p = z[mm];
            const float2 q = z[mm == 0 ? 0 : N - mm];
            // f_N = (Z_m + conj Z_{N-m})/2, f_S = (Z_m - conj Z_{N-m})/(2i)
            const float h = C.adj ? (mm == 0 || 2 * mm == N ? 0.5f * wj : wj) : 0.5f;
            const float nre = h * (p.x + q.x), nim = h * (p.y - q.y);
            const float sre = h * (p.y + q.y), sim = h * (q.x - p.x);
            gstore(fb + (unsigned)(mm * K + col), make_float4(nre + sre, nim + sim, nre - sre, nim - sim));
        }
    }
}

// ================================================================== entry points

const void *fft_kernel(int stage, bool batch, bool accumulate) {
    if (stage == kFFTSynthesis) {
        if (accumulate && !batch) return reinterpret_cast<const void *>(&fft_synthesis_kernel<false, true>);
        return batch ? reinterpret_cast<const void *>(&fft_synthesis_kernel<true>)
                     : reinterpret_cast<const void *>(&fft_synthesis_kernel<false>);
    }
    return batch ? reinterpret_cast<const void *>(&fft_analysis_kernel<true>)
                 : reinterpret_cast<const void *>(&fft_analysis_kernel<false>);
}

}  // namespace sht
