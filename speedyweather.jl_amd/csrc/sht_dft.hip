// gfx950 ring DFT of the spherical harmonic transform (Fourier stage) on MFMA.
//
// Every ring of the supported grids has N = 4p points (octahedral nlon = 16 + 4j, full
// Gaussian nlon = 4 nlat_half).  As in the LDS FFT, one complex transform of length N per
// ring and column carries the northern ring in the real and the mirrored southern ring in
// the imaginary part (the per-ring rfft / brfft pairs of SpeedyTransforms/src/fourier.jl:
// 27-77,129-250 for both hemispheres at once).  The length-N DFT splits as 4 x p:
//   analysis   F[4 m1 + m2] = sum_n1 w_p^{m1 n1} X[n1, m2],
//              X[n1, m2]    = w_N^{m2 n1} sum_n2 z[n1 + p n2] w_4^{m2 n2}
//   synthesis  z[n1 + p n2] = sum_m2 w_4^{-m2 n2} w_N^{-m2 n1} Y[n1, m2],
//              Y[n1, m2]    = sum_m1 w_p^{-m1 n1} Z[4 m1 + m2]
// (w_L = exp(-2 pi i / L)).  The radix-4 step runs on the vector units; the p-point DFT
// runs as two real GEMMs on MFMA over the symmetric pairs S_t = X[t] + X[p-t],
// D_t = X[t] - X[p-t] (t = 1..(p-1)/2):
//   T1[r] = sum_t cos(2 pi r t / p) S_t,   T2[r] = sum_t sin(2 pi r t / p) D_t,
//   sum_n1 w_p^{+-r n1} X[n1] = X[0] + (-1)^r X[p/2] + T1[r] -+ i T2[r]   (p/2 term if p even)
// and row r of T1, T2 gives both outputs r and p - r.  GEMM rows r = 0..p/2 (tiles of 16),
// contraction t, columns (column k, m2, re/im) = 8 per data column.  Only the outputs the
// Legendre stage uses are computed: an analysis stops at the rows its orders m <= mt need,
// a synthesis contracts only over the nonzero input orders.
//
// Arithmetic: FP32 products (v_mfma_f32_16x16x4_f32) or, in the split modes, scaled FP16
// hi/lo parts of both operands with three products (v_mfma_f32_16x16x32_f16), FP32
// accumulation either way.  The cos / sin matrices come precomputed from FP64 in fragment
// order; the data operand is staged in LDS (columns contiguous in t).
#include "sht_device.h"

namespace sht {

namespace {

constexpr int kMaxP = kFFTMaxRing / 4;

// Ring and column group of a workgroup (CG columns per group).  Single calls: blocks
// (ring rank, column group), rings heaviest first; batches: from the tile list.
struct DftTile {
    int j, N, p, mt, k0, ncol;
};

template <bool BATCH, int CG>
__device__ inline bool dft_tile(const DevPlan &P, const Tile &T, int K, DftTile &t) {
    int cg;
    if (BATCH) {
        if (T.m < 0) return false;
        t.j = T.m;
        cg = T.blk;
    } else {
        const int kgs = (K + CG - 1) / CG;
        const int b = blockIdx.x;
        const int r = b / kgs;
        cg = b - r * kgs;
        if (r >= P.n) return false;
        t.j = as_global(P.ring_order)[r];
    }
    t.N = as_global(P.nlon)[t.j];
    t.p = t.N >> 2;
    t.mt = as_global(P.mt)[t.j];
    t.k0 = cg * CG;
    t.ncol = min(CG, K - t.k0);
    return t.ncol > 0;
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Stages one value pair (S, D) of column c, contraction index k (0-based) into the B
// operand: FP32 as is, FP16 as hi / lo parts of the value times `s`.
template <bool F16>
__device__ __forceinline__ void stage_sd(unsigned char *sm, int bs, int cols, int c, int k, float sv, float dv,
                                         float s) {
    if constexpr (!F16) {
        float *B = reinterpret_cast<float *>(sm);
        B[c * bs + k] = sv;
        B[(cols + c) * bs + k] = dv;
    } else {
        u16 *B = reinterpret_cast<u16 *>(sm);
        const float a = sv * s, b = dv * s;
        const _Float16 ah = (_Float16)a, bh = (_Float16)b;
        const _Float16 al = (_Float16)(a - (float)ah), bl = (_Float16)(b - (float)bh);
        B[c * bs + k] = __builtin_bit_cast(u16, ah);
        B[(cols + c) * bs + k] = __builtin_bit_cast(u16, al);
        B[(2 * cols + c) * bs + k] = __builtin_bit_cast(u16, bh);
        B[(3 * cols + c) * bs + k] = __builtin_bit_cast(u16, bl);
    }
}

// The two GEMMs T1 = C S, T2 = Sn D for the row tiles and column tiles of this wave.
// Work units are (row tile, group of NT/ng column tiles); wave w takes units w, w+4, w+8.
// acc[u][ct][0] = T1, [1] = T2 of unit u's column tile ct.
constexpr int kDftUnits = 3;
template <bool F16, int CG>
struct DftAcc {
    static constexpr int NT = CG / 2;  // column tiles (16 GEMM columns each)
    f4 a[kDftUnits][NT][2];
};

// KC chunks are contracted; a row tile of the ring's table holds KCt >= KC chunks.
template <bool F16, int CG>
__device__ inline void dft_gemm(const DevPlan &P, int j, const unsigned char *sm, int bs, int Mt, int KC, int KCt,
                                int wave, int lane, DftAcc<F16, CG> &A, int &ng) {
    constexpr int NT = CG / 2;
    constexpr int cols = 8 * CG;
    ng = Mt >= 4 ? 1 : (Mt == 1 ? NT : (NT >= 2 ? 2 : 1));  // column groups per row tile
    const int ntg = NT / ng;                                // column tiles per unit
    const int row = lane & 15, grp = lane >> 4;
#pragma unroll
    for (int u = 0; u < kDftUnits; ++u)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) A.a[u][ct][0] = A.a[u][ct][1] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < kDftUnits; ++u) {
        const int unit = wave + 4 * u;
        if (unit >= Mt * ng) break;  // uniform per wave
        const int rt = unit / ng, g0 = (unit - rt * ng) * ntg;
        if constexpr (!F16) {
            const SHT_GLOBAL float4 *At = reinterpret_cast<const SHT_GLOBAL float4 *>(
                                              as_global(P.dft32) + as_global(P.dft32_off)[j]) +
                                          ((long long)rt * KCt * 64 + lane) * 2;
            const float *Bs = reinterpret_cast<const float *>(sm);
            float4 ac = gload(At), as = gload(At + 1);
            for (int ch = 0; ch < KC; ++ch) {
                const float4 c4 = ac, s4 = as;
                const int nx = min(ch + 1, KC - 1);
                ac = gload(At + (long long)nx * 128);
                as = gload(At + (long long)nx * 128 + 1);
#pragma unroll
                for (int ct = 0; ct < NT; ++ct) {
                    if (ct >= ntg) break;
                    const int c = (g0 + ct) * 16 + row;
                    const float4 bsv = *reinterpret_cast<const float4 *>(Bs + c * bs + ch * kDftChunk32 + 4 * grp);
                    const float4 bdv =
                        *reinterpret_cast<const float4 *>(Bs + (cols + c) * bs + ch * kDftChunk32 + 4 * grp);
                    f4 t1 = A.a[u][ct][0], t2 = A.a[u][ct][1];
                    t1 = mfma_f32(c4.x, bsv.x, t1);
                    t2 = mfma_f32(s4.x, bdv.x, t2);
                    t1 = mfma_f32(c4.y, bsv.y, t1);
                    t2 = mfma_f32(s4.y, bdv.y, t2);
                    t1 = mfma_f32(c4.z, bsv.z, t1);
                    t2 = mfma_f32(s4.z, bdv.z, t2);
                    t1 = mfma_f32(c4.w, bsv.w, t1);
                    t2 = mfma_f32(s4.w, bdv.w, t2);
                    A.a[u][ct][0] = t1;
                    A.a[u][ct][1] = t2;
                }
            }
        } else {
            typedef Split<kFP16x3>::V V8;
            const SHT_GLOBAL V8 *At = reinterpret_cast<const SHT_GLOBAL V8 *>(as_global(P.dft16) +
                                                                             as_global(P.dft16_off)[j]) +
                                      ((long long)rt * KCt * 64 + lane) * 4;
            const u16 *B = reinterpret_cast<const u16 *>(sm);
            V8 ch_ = At[0], cl_ = At[1], sh_ = At[2], sl_ = At[3];
            for (int ch = 0; ch < KC; ++ch) {
                const V8 chi = ch_, clo = cl_, shi = sh_, slo = sl_;
                const long long nx = (long long)min(ch + 1, KC - 1) * 256;
                ch_ = At[nx];
                cl_ = At[nx + 1];
                sh_ = At[nx + 2];
                sl_ = At[nx + 3];
#pragma unroll
                for (int ct = 0; ct < NT; ++ct) {
                    if (ct >= ntg) break;
                    const int c = (g0 + ct) * 16 + row;
                    const int o = ch * kDftChunk16 + 8 * grp;
                    const V8 bsh = __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(B + c * bs + o));
                    const V8 bsl = __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(B + (cols + c) * bs + o));
                    const V8 bdh =
                        __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(B + (2 * cols + c) * bs + o));
                    const V8 bdl =
                        __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(B + (3 * cols + c) * bs + o));
                    f4 t1 = A.a[u][ct][0], t2 = A.a[u][ct][1];
                    t1 = Split<kFP16x3>::mma(clo, bsh, t1);
                    t2 = Split<kFP16x3>::mma(slo, bdh, t2);
                    t1 = Split<kFP16x3>::mma(chi, bsl, t1);
                    t2 = Split<kFP16x3>::mma(shi, bdl, t2);
                    t1 = Split<kFP16x3>::mma(chi, bsh, t1);
                    t2 = Split<kFP16x3>::mma(shi, bdh, t2);
                    A.a[u][ct][0] = t1;
                    A.a[u][ct][1] = t2;
                }
            }
        }
    }
}

// Power-of-two FP16 scale of column group kc from the LDS maximum (bits of a float >= 0).
__device__ __forceinline__ float dft_scale_of(const unsigned *amax_bits, int kc, float &inv) {
    float s, si;
    fp16_scale(__uint_as_float(amax_bits[kc]), s, si);
    inv = si;
    return s;
}

}  // namespace

// =================================================================== analysis
//
// grid -> Fourier rows (fourier.jl:129-166): per ring and column, F = DFT(f_N + i f_S),
// separated by symmetry into f_N, f_S and stored as (f_N + f_S, f_N - f_S) for m <= mt.
template <bool F16, int CG, bool BATCH>
__global__ __launch_bounds__(256) void dft_analysis_kernel(const DevPlan *__restrict__ plans, CallArgs single,
                                                           const CallArgs *__restrict__ calls,
                                                           const Tile *__restrict__ tiles) {
    extern __shared__ float4 dsm4[];
    unsigned char *sm = reinterpret_cast<unsigned char *>(dsm4);
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Tile T = BATCH ? tiles[blockIdx.x] : Tile{0, 0, 0, 0};
    const CallArgs C = BATCH ? calls[T.call] : single;
    const DevPlan &P = plans[C.plan];
    const int K = C.K;
    DftTile G;
    if (!dft_tile<BATCH, CG>(P, T, K, G)) return;  // padding of the launch order
    const int N = G.N, p = G.p, mt = G.mt, ncol = G.ncol;
    const int h = (p - 1) >> 1, half = p >> 1;
    const bool even = (p & 1) == 0;
    constexpr int cols = 8 * CG;
    const int KC = (h + (F16 ? kDftChunk16 : kDftChunk32) - 1) / (F16 ? kDftChunk16 : kDftChunk32);
    const int kp = dft_kpad(h < 1 ? 1 : h, F16), bs = dft_bstride(kp, F16);
    const int bbytes = 2 * cols * bs * 4;
    const int FS = 2 * mt + 1;  // staged outputs per column: F[0..mt], F[N-1..N-mt]
    const int fbytes = CG * (FS + 1) * 8;
    unsigned char *aux = sm + (bbytes > fbytes ? bbytes : fbytes);
    float2 *X0 = reinterpret_cast<float2 *>(aux);  // [kc][m2]
    float2 *Xh = X0 + CG * 4;
    unsigned *amax = reinterpret_cast<unsigned *>(Xh + CG * 4);
    if (F16 && tid < CG) amax[tid] = 0u;

    const SHT_GLOBAL float *field = as_global(static_cast<const float *>(C.in));
    const long long ld = C.ld_in;
    const long long rn = as_global(P.ring_n)[G.j], rs = as_global(P.ring_s)[G.j];
    const SHT_GLOBAL float2 *tw = as_global(reinterpret_cast<const float2 *>(P.twiddle)) + as_global(P.tw_off)[G.j];

    // items (t, kc), t = 0..p/2 fastest: X[t] and X[p - t] of column kc
    const int NI = half + 1;
    constexpr int U = (CG * (kMaxP / 2 + 1) + 255) / 256;
    float sd[U][16];  // S then D of (m2, re/im), or X0 / Xh
    float lmax[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int it = tid + 256 * u;
        const int t = it % NI, kc = it / NI;
        if (kc >= CG) break;
        const int kq = G.k0 + min(kc, ncol - 1);  // absent columns re-read the last one, not stored
        const SHT_GLOBAL float *fn = field + (long long)kq * ld + rn;
        const SHT_GLOBAL float *fs = field + (long long)kq * ld + rs;
        const int tb = t == 0 ? 0 : p - t;
        float2 a[4], b[4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            a[n2] = make_float2(fn[t + p * n2], fs[t + p * n2]);
            b[n2] = make_float2(fn[tb + p * n2], fs[tb + p * n2]);
        }
        // radix 4 (forward: w_4 = -i) and twiddles w_N^{m2 n1}
        float2 xa[4], xb[4];
        {
            const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
            const float2 s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
            xa[0] = cadd(s02, s13);
            xa[1] = make_float2(d02.x + d13.y, d02.y - d13.x);  // d02 - i d13
            xa[2] = csub(s02, s13);
            xa[3] = make_float2(d02.x - d13.y, d02.y + d13.x);  // d02 + i d13
        }
        {
            const float2 s02 = cadd(b[0], b[2]), d02 = csub(b[0], b[2]);
            const float2 s13 = cadd(b[1], b[3]), d13 = csub(b[1], b[3]);
            xb[0] = cadd(s02, s13);
            xb[1] = make_float2(d02.x + d13.y, d02.y - d13.x);
            xb[2] = csub(s02, s13);
            xb[3] = make_float2(d02.x - d13.y, d02.y + d13.x);
        }
#pragma unroll
        for (int m2 = 1; m2 < 4; ++m2) {
            xa[m2] = cmulf(xa[m2], gload(tw + m2 * t));
            xb[m2] = cmulf(xb[m2], gload(tw + m2 * tb));
        }
        if (t == 0 || (even && t == half)) {
            float2 *dst = (t == 0 ? X0 : Xh) + kc * 4;
#pragma unroll
            for (int m2 = 0; m2 < 4; ++m2) dst[m2] = xa[m2];
        }
#pragma unroll
        for (int m2 = 0; m2 < 4; ++m2) {
            sd[u][4 * m2 + 0] = xa[m2].x + xb[m2].x;
            sd[u][4 * m2 + 1] = xa[m2].y + xb[m2].y;
            sd[u][4 * m2 + 2] = xa[m2].x - xb[m2].x;
            sd[u][4 * m2 + 3] = xa[m2].y - xb[m2].y;
        }
        lmax[u] = 0.f;
        if (F16 && t >= 1 && t <= h)
#pragma unroll
            for (int e = 0; e < 16; ++e) lmax[u] = fmaxf(lmax[u], fabsf(sd[u][e]));
    }
    float scl[U];
#pragma unroll
    for (int u = 0; u < U; ++u) scl[u] = 1.f;
    if constexpr (F16) {
        __syncthreads();  // amax initialised
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int it = tid + 256 * u;
            const int kc = it / NI;
            if (kc >= CG) break;
            const int t = it % NI;
            if (t >= 1 && t <= h) atomicMax(amax + kc, __float_as_uint(lmax[u]));
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kc = min((tid + 256 * u) / NI, CG - 1);
            float inv;
            scl[u] = dft_scale_of(amax, kc, inv);
        }
    }
    // stage S, D (t = 1..h at k = t - 1) and zero the padding k = h..kp-1
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int it = tid + 256 * u;
        const int t = it % NI, kc = it / NI;
        if (kc >= CG) break;
        if (t < 1 || t > h) continue;
#pragma unroll
        for (int m2 = 0; m2 < 4; ++m2)
#pragma unroll
            for (int ri = 0; ri < 2; ++ri)
                stage_sd<F16>(sm, bs, cols, kc * 8 + m2 * 2 + ri, t - 1, sd[u][4 * m2 + ri], sd[u][4 * m2 + 2 + ri],
                              scl[u]);
    }
    for (int i = tid; i < cols * (kp - h); i += 256) {
        const int c = i % cols, k = h + i / cols;
        stage_sd<F16>(sm, bs, cols, c, k, 0.f, 0.f, 1.f);
    }
    __syncthreads();

    const int rows = dft_rows_analysis(p, mt);
    const int Mt = (rows + 15) >> 4;
    DftAcc<F16, CG> acc;
    int ng;
    dft_gemm<F16, CG>(P, G.j, sm, bs, Mt, KC, KC, wave, lane, acc, ng);
    __syncthreads();  // the staged outputs overwrite the operands

    float2 *Fs = reinterpret_cast<float2 *>(sm);
    {
        constexpr int NT = CG / 2;
        const int ntg = NT / ng;
        const int row = lane & 15, grp = lane >> 4;
#pragma unroll
        for (int u = 0; u < kDftUnits; ++u) {
            const int unit = wave + 4 * u;
            if (unit >= Mt * ng) break;
            const int rt = unit / ng, g0 = (unit - rt * ng) * ntg;
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                if (ct >= ntg) break;
                const int c = (g0 + ct) * 16 + row;
                const int kc = c >> 3, m2 = (c >> 1) & 3, reim = c & 1;
                float sc = 1.f;
                if constexpr (F16) {
                    float inv;
                    dft_scale_of(amax, kc, inv);
                    sc = inv * (1.f / kDftAScale);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float v1 = acc.a[u][ct][0][i] * sc, v2 = acc.a[u][ct][1][i] * sc;
                    const float p1 = lane_xor1(v1), p2 = lane_xor1(v2);
                    const int r = rt * 16 + 4 * grp + i;
                    if (r >= rows || r > half) continue;
                    const float2 T1 = reim ? make_float2(p1, v1) : make_float2(v1, p1);
                    const float2 T2 = reim ? make_float2(p2, v2) : make_float2(v2, p2);
                    float2 b = X0[kc * 4 + m2];
                    if (even) {
                        const float2 xh = Xh[kc * 4 + m2];
                        const float sg = (r & 1) ? -1.f : 1.f;
                        b = make_float2(b.x + sg * xh.x, b.y + sg * xh.y);
                    }
                    // lane re: output r = base + T1 - i T2; lane im: output p - r = base + T1 + i T2
                    int q;
                    float2 g;
                    if (reim == 0) {
                        q = r;
                        g = make_float2(b.x + T1.x + T2.y, b.y + T1.y - T2.x);
                    } else {
                        if (r == 0 || (even && r == half)) continue;
                        q = p - r;
                        g = make_float2(b.x + T1.x - T2.y, b.y + T1.y + T2.x);
                    }
                    const int m = 4 * q + m2;
                    if (m <= mt) Fs[kc * FS + m] = g;
                    if (m > 0 && m >= N - mt) Fs[kc * FS + mt + (N - m)] = g;
                }
            }
        }
    }
    __syncthreads();
    {
        SHT_GLOBAL float4 *fb = as_global(static_cast<float4 *>(C.fbuf)) + as_global(P.fboff)[G.j] * (long long)K + G.k0;
        // adjoint of the synthesis (spec2grid pullback): rows times s_m (1/cos lat) / ΔΩ_j
        const float wj = C.adj ? (C.unscale ? as_global(P.coslat_inv)[G.j] : 1.f) / as_global(P.domega)[G.j] : 1.f;
        // rows mt < m <= mte: zeros (a non-monotonic shortcut's orders of rings nearer the pole)
        const int mte = as_global(P.mte)[G.j];
        for (int i = tid; i < CG * (mte + 1); i += 256) {
            const int kc = i % CG, mm = i / CG;
            if (kc >= ncol) continue;
            if (mm > mt) {
                gstore(fb + (unsigned)(mm * K + kc), make_float4(0.f, 0.f, 0.f, 0.f));
                continue;
            }
            const float2 pv = Fs[kc * FS + mm];
            const float2 qv = mm == 0 ? pv : Fs[kc * FS + mt + mm];
            // f_N = (Z_m + conj Z_{N-m})/2, f_S = (Z_m - conj Z_{N-m})/(2i)
            const float hh = C.adj ? (mm == 0 || 2 * mm == N ? 0.5f * wj : wj) : 0.5f;
            const float nre = hh * (pv.x + qv.x), nim = hh * (pv.y - qv.y);
            const float sre = hh * (pv.y + qv.y), sim = hh * (qv.x - pv.x);
            gstore(fb + (unsigned)(mm * K + kc), make_float4(nre + sre, nim + sim, nre - sre, nim - sim));
        }
    }
}

// ================================================================== synthesis
//
// Fourier rows -> grid (fourier.jl:215-250): per ring and column, the inverse DFT of the
// Hermitian-extended Z = g_N + i g_S (Im of the DC and Nyquist bins ignored, as FFTW's c2r
// does; orders above mt are zero); Re -> northern ring, Im -> southern ring.  ACC: adds
// into the grid (pullbacks).
template <bool F16, int CG, bool BATCH, bool ACC>
__global__ __launch_bounds__(256) void dft_synthesis_kernel(const DevPlan *__restrict__ plans, CallArgs single,
                                                            const CallArgs *__restrict__ calls,
                                                            const Tile *__restrict__ tiles) {
    extern __shared__ float4 dsm4[];
    unsigned char *sm = reinterpret_cast<unsigned char *>(dsm4);
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Tile T = BATCH ? tiles[blockIdx.x] : Tile{0, 0, 0, 0};
    const CallArgs C = BATCH ? calls[T.call] : single;
    const DevPlan &P = plans[C.plan];
    const int K = C.K;
    DftTile G;
    if (!dft_tile<BATCH, CG>(P, T, K, G)) return;
    const int N = G.N, p = G.p, mt = G.mt, ncol = G.ncol;
    const int h = (p - 1) >> 1, half = p >> 1;
    const bool even = (p & 1) == 0;
    constexpr int cols = 8 * CG;
    const int Kd = dft_k_synthesis(p, mt);
    const int CH = F16 ? kDftChunk16 : kDftChunk32;
    const int KC = (Kd + CH - 1) / CH;
    const int kp = dft_kpad(Kd < 1 ? 1 : Kd, F16), bs = dft_bstride(kp, F16);
    const int bbytes = 2 * cols * bs * 4;
    const int zbytes = CG * (mt + 1) * 16;
    const int ybytes = CG * p * 4 * 8;
    unsigned char *aux = sm + (zbytes + bbytes > ybytes ? zbytes + bbytes : ybytes);
    float2 *X0 = reinterpret_cast<float2 *>(aux);
    float2 *Xh = X0 + CG * 4;
    unsigned *amax = reinterpret_cast<unsigned *>(Xh + CG * 4);
    float4 *Zs = reinterpret_cast<float4 *>(sm + bbytes);  // [kc][m], after the operands
    if (F16 && tid < CG) amax[tid] = 0u;

    // Fourier rows m <= mt of the group's columns into LDS
    {
        const SHT_GLOBAL float4 *fb =
            as_global(static_cast<const float4 *>(C.fbuf)) + as_global(P.fboff)[G.j] * (long long)K + G.k0;
        const float wj = C.adj ? as_global(P.domega)[G.j] : 1.f;
        constexpr int UR = 4;  // rounds of loads in flight together
        for (int base = 0; base < CG * (mt + 1); base += 256 * UR) {
            float4 v[UR];
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                const int i = base + tid + 256 * u;
                const int kc = i % CG, mm = i / CG;
                v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (mm <= mt && kc < ncol) v[u] = gload(fb + (unsigned)(mm * K + kc));
            }
#pragma unroll
            for (int u = 0; u < UR; ++u) {
                const int i = base + tid + 256 * u;
                const int kc = i % CG, mm = i / CG;
                if (mm > mt) break;
                float4 w = v[u];
                if (C.adj) {  // adjoint of the analysis (grid2spec pullback): rows times ΔΩ_j / s_m
                    const float sw = mm == 0 || 2 * mm == N ? wj : 0.5f * wj;
                    w = make_float4(w.x * sw, w.y * sw, w.z * sw, w.w * sw);
                }
                Zs[kc * (mt + 1) + mm] = w;
            }
        }
    }
    __syncthreads();
    // Z at index q (0 <= q < N) from the staged rows: forward form of row q (q <= N/2) or
    // mirrored form of row N - q; zero beyond mt.
    auto Zat = [&](int kc, int q) -> float2 {
        const bool fwd = 2 * q <= N;
        const int mm = fwd ? q : N - q;
        if (mm > mt) return make_float2(0.f, 0.f);
        const float4 w = Zs[kc * (mt + 1) + mm];
        if (mm == 0 || 2 * mm == N) return make_float2(w.x, w.z);
        return fwd ? make_float2(w.x - w.w, w.y + w.z) : make_float2(w.x + w.w, w.z - w.y);
    };
    // items (t, kc), t = 0..Kd, then the p/2 item: S_t, D_t of W[t] = Z[4t + m2], W[p - t]
    const int NI = Kd + 2;
    constexpr int U = (CG * (kMaxP / 2 + 2) + 255) / 256;
    float sd[U][16];
    float lmax[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int it = tid + 256 * u;
        const int t = it % NI, kc = it / NI;
        lmax[u] = 0.f;
        if (kc >= CG) break;
        if (t == Kd + 1) {  // W[p/2][m2] = Z[2p + m2]
            if (even)
#pragma unroll
                for (int m2 = 0; m2 < 4; ++m2) Xh[kc * 4 + m2] = Zat(kc, 2 * p + m2);
            continue;
        }
#pragma unroll
        for (int m2 = 0; m2 < 4; ++m2) {
            const float2 wa = Zat(kc, 4 * t + m2);
            if (t == 0) {
                X0[kc * 4 + m2] = wa;
                continue;
            }
            const float2 wb = Zat(kc, 4 * (p - t) + m2);
            sd[u][4 * m2 + 0] = wa.x + wb.x;
            sd[u][4 * m2 + 1] = wa.y + wb.y;
            sd[u][4 * m2 + 2] = wa.x - wb.x;
            sd[u][4 * m2 + 3] = wa.y - wb.y;
            if (F16)
#pragma unroll
                for (int e = 0; e < 4; ++e) lmax[u] = fmaxf(lmax[u], fabsf(sd[u][4 * m2 + e]));
        }
    }
    float scl[U];
#pragma unroll
    for (int u = 0; u < U; ++u) scl[u] = 1.f;
    if constexpr (F16) {
        // every thread reached here after the staged-row barrier, so amax is initialised
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int it = tid + 256 * u;
            const int kc = it / NI, t = it % NI;
            if (kc >= CG) break;
            if (t >= 1 && t <= Kd) atomicMax(amax + kc, __float_as_uint(lmax[u]));
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kc = min((tid + 256 * u) / NI, CG - 1);
            float inv;
            scl[u] = dft_scale_of(amax, kc, inv);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int it = tid + 256 * u;
        const int t = it % NI, kc = it / NI;
        if (kc >= CG) break;
        if (t < 1 || t > Kd) continue;
#pragma unroll
        for (int m2 = 0; m2 < 4; ++m2)
#pragma unroll
            for (int ri = 0; ri < 2; ++ri)
                stage_sd<F16>(sm, bs, cols, kc * 8 + m2 * 2 + ri, t - 1, sd[u][4 * m2 + ri], sd[u][4 * m2 + 2 + ri],
                              scl[u]);
    }
    for (int i = tid; i < cols * (kp - Kd); i += 256) {
        const int c = i % cols, k = Kd + i / cols;
        stage_sd<F16>(sm, bs, cols, c, k, 0.f, 0.f, 1.f);
    }
    __syncthreads();

    const int Mt = (half + 1 + 15) >> 4;
    DftAcc<F16, CG> acc;
    int ng;
    const int KCt = (h + CH - 1) / CH;
    dft_gemm<F16, CG>(P, G.j, sm, bs, Mt, KC, KCt, wave, lane, acc, ng);
    __syncthreads();  // Y overwrites the operands and the staged rows

    float2 *Ys = reinterpret_cast<float2 *>(sm);  // [kc][n1][m2]
    {
        constexpr int NT = CG / 2;
        const int ntg = NT / ng;
        const int row = lane & 15, grp = lane >> 4;
#pragma unroll
        for (int u = 0; u < kDftUnits; ++u) {
            const int unit = wave + 4 * u;
            if (unit >= Mt * ng) break;
            const int rt = unit / ng, g0 = (unit - rt * ng) * ntg;
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                if (ct >= ntg) break;
                const int c = (g0 + ct) * 16 + row;
                const int kc = c >> 3, m2 = (c >> 1) & 3, reim = c & 1;
                float sc = 1.f;
                if constexpr (F16) {
                    float inv;
                    dft_scale_of(amax, kc, inv);
                    sc = inv * (1.f / kDftAScale);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float v1 = acc.a[u][ct][0][i] * sc, v2 = acc.a[u][ct][1][i] * sc;
                    const float p1 = lane_xor1(v1), p2 = lane_xor1(v2);
                    const int r = rt * 16 + 4 * grp + i;
                    if (r > half) continue;
                    const float2 T1 = reim ? make_float2(p1, v1) : make_float2(v1, p1);
                    const float2 T2 = reim ? make_float2(p2, v2) : make_float2(v2, p2);
                    float2 b = X0[kc * 4 + m2];
                    if (even) {
                        const float2 xh = Xh[kc * 4 + m2];
                        const float sg = (r & 1) ? -1.f : 1.f;
                        b = make_float2(b.x + sg * xh.x, b.y + sg * xh.y);
                    }
                    // lane re: Y[r] = base + T1 + i T2; lane im: Y[p - r] = base + T1 - i T2
                    if (reim == 0) {
                        Ys[(kc * p + r) * 4 + m2] = make_float2(b.x + T1.x - T2.y, b.y + T1.y + T2.x);
                    } else if (r != 0 && !(even && r == half)) {
                        Ys[(kc * p + p - r) * 4 + m2] = make_float2(b.x + T1.x + T2.y, b.y + T1.y - T2.x);
                    }
                }
            }
        }
    }
    __syncthreads();
    {
        SHT_GLOBAL float *field = as_global(static_cast<float *>(C.out));
        const long long ld = C.ld_out;
        const long long rn = as_global(P.ring_n)[G.j], rs = as_global(P.ring_s)[G.j];
        const SHT_GLOBAL float2 *tw =
            as_global(reinterpret_cast<const float2 *>(P.twiddle)) + as_global(P.tw_off)[G.j];
        for (int i = tid; i < ncol * p; i += 256) {
            const int n1 = i % p, kc = i / p;
            float2 y[4];
#pragma unroll
            for (int m2 = 0; m2 < 4; ++m2) y[m2] = Ys[(kc * p + n1) * 4 + m2];
#pragma unroll
            for (int m2 = 1; m2 < 4; ++m2) y[m2] = cmulc(y[m2], gload(tw + m2 * n1));  // w_N^{-m2 n1}
            // inverse radix 4: z[n1 + p n2] = sum_m2 i^{m2 n2} y[m2]
            const float2 s02 = cadd(y[0], y[2]), d02 = csub(y[0], y[2]);
            const float2 s13 = cadd(y[1], y[3]), d13 = csub(y[1], y[3]);
            float2 z[4];
            z[0] = cadd(s02, s13);
            z[1] = make_float2(d02.x - d13.y, d02.y + d13.x);  // d02 + i d13
            z[2] = csub(s02, s13);
            z[3] = make_float2(d02.x + d13.y, d02.y - d13.x);  // d02 - i d13
            SHT_GLOBAL float *fn = field + (long long)(G.k0 + kc) * ld + rn;
            SHT_GLOBAL float *fs = field + (long long)(G.k0 + kc) * ld + rs;
            if constexpr (ACC) {
                float on[4], os[4];
#pragma unroll
                for (int n2 = 0; n2 < 4; ++n2) {
                    on[n2] = fn[n1 + p * n2];
                    os[n2] = fs[n1 + p * n2];
                }
#pragma unroll
                for (int n2 = 0; n2 < 4; ++n2) {
                    fn[n1 + p * n2] = on[n2] + z[n2].x;
                    fs[n1 + p * n2] = os[n2] + z[n2].y;
                }
            } else {
#pragma unroll
                for (int n2 = 0; n2 < 4; ++n2) {
                    fn[n1 + p * n2] = z[n2].x;
                    fs[n1 + p * n2] = z[n2].y;
                }
            }
        }
    }
}

// ================================================================== entry points

template <bool F16, int CG>
static const void *dft_entry(int stage, bool batch, bool accumulate) {
    if (stage == kFFTSynthesis) {
        if (accumulate && !batch) return reinterpret_cast<const void *>(&dft_synthesis_kernel<F16, CG, false, true>);
        return batch ? reinterpret_cast<const void *>(&dft_synthesis_kernel<F16, CG, true, false>)
                     : reinterpret_cast<const void *>(&dft_synthesis_kernel<F16, CG, false, false>);
    }
    return batch ? reinterpret_cast<const void *>(&dft_analysis_kernel<F16, CG, true>)
                 : reinterpret_cast<const void *>(&dft_analysis_kernel<F16, CG, false>);
}

const void *dft_kernel(int stage, bool f16, int cg, bool batch, bool accumulate) {
    if (f16) return cg == 8 ? dft_entry<true, 8>(stage, batch, accumulate) : dft_entry<true, 4>(stage, batch, accumulate);
    return cg == 8 ? dft_entry<false, 8>(stage, batch, accumulate) : dft_entry<false, 4>(stage, batch, accumulate);
}

}  // namespace sht
