// gfx950 Legendre-stage kernels of the spherical harmonic transform.
//
// For every order m a dense contraction between the precomputed Legendre polynomials
// Λ_m (degrees x rings) and the K spectral / Fourier columns, split into even and odd
// degrees l-m so one contraction serves both hemispheres (the even/odd symmetry of
// SpeedyTransforms/src/legendre.jl:2-33,104-126).  Runs on MFMA: v_mfma_f32_16x16x4_f32
// in FP32 mode; v_mfma_f32_16x16x32_{f16,bf16} in the split modes, FP32 accumulation:
// scaled FP16 hi/lo parts of both operands (three products), or BF16 hi/mid/lo parts of
// both operands (six products, every term down to 2^-16 of the leading one).
//
// The contraction dimension is streamed in chunks of 32 (degrees per parity, or rings):
// the next chunk's operands are loaded into registers while the current chunk's MFMAs
// run, and the staged operand tile in LDS is double-buffered (one barrier per chunk).
//
// Fourier buffer (the scratch of the call's stream): row (fboff[j] + m) holds, for every column k, a
// float4 (N.re, N.im, S.re, S.im).  Synthesis writes the northern/southern Legendre sums
// g_N, g_S there (legendre.jl:84-91); analysis reads the even/odd combinations
// f_N + f_S, f_N - f_S (legendre.jl:176-183, the solid angle folded into Λ).
#include "sht_device.h"

namespace sht {

typedef u16 u16x4 __attribute__((ext_vector_type(4)));

// LDS tiles of one chunk: [buffer][parity][real column][contraction index], padded rows.
// In the split modes the padding of each hi row holds the row's inverse FP16 scale; the
// synthesis keeps per-ring epilogue factors in the padding of 64 further rows (ep_slot).
constexpr int kRowF = kChunk + 4;  // FP32 row (floats): 16-byte aligned, bank-spread
constexpr int kRowH = kChunk + 8;  // 16-bit row: 16-byte aligned, bank-spread
constexpr int kTileF = 2 * kColsReal * kRowF;
constexpr int kTileH = 2 * kColsReal * kRowH;

template <int MODE>
struct LegSmem {
    static constexpr int bytes = MODE == kFP32 ? 2 * kTileF * 4 : 2 * split_parts(MODE) * kTileH * 2;
};

template <int MODE>
using V8of = typename Split<MODE == kFP32 ? kFP16x3 : MODE>::V;

// Part q of a split Λ array set: hi (q = 0), then mid (BF16, q = 1), then lo (the last part).
template <int MODE>
__device__ __forceinline__ const void *lam_part(const void *hi, const void *mid, const void *lo, int q) {
    return q == 0 ? hi : (q == split_parts(MODE) - 1 ? lo : mid);
}

// Stages 4 contraction values x (re, im) of one complex column (real columns 2kk, 2kk+1)
// and parity p into the chunk tile: FP32 as is; split modes as 16-bit parts (tile q of the
// buffer holds part q): scaled FP16 hi/lo with one power-of-two scale per (parity, real
// column) over the 32 values the 8 lanes of a column group hold, or BF16 hi/mid/lo, whose
// sum is the FP32 value exactly.
template <int MODE>
__device__ inline void stage4(unsigned char *sm, int buf, int p, int kk, int g, const float (&re)[4],
                              const float (&im)[4]) {
    if constexpr (MODE == kFP32) {
        float *B = reinterpret_cast<float *>(sm) + buf * kTileF + p * kColsReal * kRowF;
        *reinterpret_cast<float4 *>(B + (2 * kk) * kRowF + 4 * g) = make_float4(re[0], re[1], re[2], re[3]);
        *reinterpret_cast<float4 *>(B + (2 * kk + 1) * kRowF + 4 * g) = make_float4(im[0], im[1], im[2], im[3]);
    } else {
        typedef typename Split<MODE>::T T16;
        constexpr int NP = split_parts(MODE);
        u16 *B0 = reinterpret_cast<u16 *>(sm) + (buf * NP) * kTileH + p * kColsReal * kRowH;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const float *x = c == 0 ? re : im;
            float s = 1.f, sinv = 1.f;
            if constexpr (MODE == kFP16x3) {
                float amax = fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3])));
                // maximum over the 8 lanes of the column group (lanes 8q .. 8q + 7)
                amax = fmaxf(amax, lane_xor1(amax));
                amax = fmaxf(amax, lane_xor2(amax));
                amax = fmaxf(amax, lane_mirror8(amax));
                fp16_scale(amax, s, sinv);
            }
            u16x4 part[NP];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float v = x[e] * s;
                const T16 h = (T16)v;
                const float r = v - (float)h;  // exact
                const T16 m = (T16)r;
                part[0][e] = __builtin_bit_cast(u16, h);
                part[1][e] = __builtin_bit_cast(u16, m);
                if constexpr (NP == 3) part[2][e] = __builtin_bit_cast(u16, (T16)(r - (float)m));
            }
            const int col = 2 * kk + c;
#pragma unroll
            for (int q = 0; q < NP; ++q) *reinterpret_cast<u16x4 *>(B0 + q * kTileH + col * kRowH + 4 * g) = part[q];
            if (MODE == kFP16x3 && g == 0) *reinterpret_cast<float *>(B0 + col * kRowH + kChunk) = sinv;
        }
    }
}

// Complex column (of the workgroup's 32) that staging thread tid stages: each group of 8
// consecutive lanes q = (tid / 8) % 8 of wave w takes one column, 8 w + f(q), with f chosen
// so that the tile rows written by one LDS write instruction fall on disjoint banks (64 banks
// of 4 bytes, 256 bytes per cycle).  Split modes: 8-byte writes of 32 lanes per cycle, four
// columns 2 f(q) whose 80-byte rows must start 64 bytes apart modulo 256: f = 0, 2, 4, 6
// for q = 0..3 and 1, 3, 5, 7 for q = 4..7.  FP32: 16-byte writes of 16 lanes per cycle, two
// columns whose 144-byte rows must start 128 bytes apart modulo 256: f = 0, 4, 1, 5, ...
template <int MODE>
__device__ __forceinline__ int stage_column(int tid) {
    const int q = (tid >> 3) & 7;
    const int f = MODE == kFP32 ? (q >> 1) + 4 * (q & 1) : 2 * (q & 3) + (q >> 2);
    return (tid >> 6) * 8 + f;
}

// Batch launches: the column map of merged column groups (passed as single.in) and the
// entry of column kk of tile T, or null / an empty entry for a tile of one call.
template <bool BATCH>
__device__ __forceinline__ const SHT_GLOBAL ColRef *merged_columns(const Tile &T, const CallArgs &single) {
    if (!BATCH || T.call >= 0) return nullptr;
    return as_global(static_cast<const ColRef *>(single.in)) + (long long)(-1 - T.call) * kColsComplex;
}
__device__ __forceinline__ ColRef column_ref(const SHT_GLOBAL ColRef *cmap, int k) {
    const SHT_GLOBAL int *q = reinterpret_cast<const SHT_GLOBAL int *>(cmap + k);
    return ColRef{q[0], q[1]};
}
// The call of a merged column (an empty one reads the group's first call: same plan and
// direction, so every index stays inside its buffers; nothing is stored for it).
__device__ __forceinline__ const SHT_GLOBAL CallArgs *call_of(const CallArgs *calls, int call, int call0) {
    return as_global(calls) + (call < 0 ? call0 : call);
}

// acc[t] += A_t (16 rows x 32 contraction, this lane's fragment of row tile t < TPW) x B
// tile columns cb*16..cb*16+15 of parity p; each B fragment is read once for all TPW row
// tiles.  FP32: af[t][8] holds the lane's A values of tile t for the 8 K-steps of 4.
// Split modes: ap[t][q] holds part q of the lane's A fragment; FP16 sums lo·hi, hi·lo,
// hi·hi per chunk and adds the chunk rescaled, BF16 accumulates the six products of parts
// (i, j) with i + j <= 2, smallest first.
template <int MODE, int NCB, int TPW>
__device__ inline void mma_chunk(const unsigned char *sm, int buf, int p, int row, int grp, const float (&af)[TPW][8],
                                 const V8of<MODE> (&ap)[TPW][3], f4 (&acc)[TPW][NCB]) {
    if constexpr (MODE == kFP32) {
        const float *B = reinterpret_cast<const float *>(sm) + buf * kTileF + p * kColsReal * kRowF;
#pragma unroll
        for (int ks = 0; ks < kChunk / 4; ++ks) {
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const float b = B[(cb * 16 + row) * kRowF + 4 * ks + grp];
#pragma unroll
                for (int t = 0; t < TPW; ++t) acc[t][cb] = mfma_f32(af[t][ks], b, acc[t][cb]);
            }
        }
    } else {
        typedef V8of<MODE> V8;
        constexpr int NP = split_parts(MODE);
        const u16 *B0 = reinterpret_cast<const u16 *>(sm) + (buf * NP) * kTileH + p * kColsReal * kRowH;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const int col = cb * 16 + row;
            V8 b[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q)
                b[q] = __builtin_bit_cast(V8, *reinterpret_cast<const u16x8 *>(B0 + q * kTileH + col * kRowH + 8 * grp));
            if constexpr (MODE == kFP16x3) {
                const float sc = *reinterpret_cast<const float *>(B0 + col * kRowH + kChunk);
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    f4 u = Split<MODE>::mma(ap[t][1], b[0], f4{0.f, 0.f, 0.f, 0.f});
                    u = Split<MODE>::mma(ap[t][0], b[1], u);
                    u = Split<MODE>::mma(ap[t][0], b[0], u);
                    acc[t][cb] += u * sc;
                }
            } else {
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    f4 c = acc[t][cb];
                    c = Split<MODE>::mma(ap[t][2], b[0], c);
                    c = Split<MODE>::mma(ap[t][0], b[2], c);
                    c = Split<MODE>::mma(ap[t][1], b[1], c);
                    c = Split<MODE>::mma(ap[t][1], b[0], c);
                    c = Split<MODE>::mma(ap[t][0], b[1], c);
                    acc[t][cb] = Split<MODE>::mma(ap[t][0], b[0], c);
                }
            }
        }
    }
}

// The same for one row tile.
template <int MODE, int NCB>
__device__ inline void mma_chunk1(const unsigned char *sm, int buf, int p, int row, int grp, const float (&af)[8],
                                  const V8of<MODE> (&ap)[3], f4 (&acc)[NCB]) {
    float a1[1][8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) a1[0][ks] = af[ks];
    V8of<MODE> p1[1][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) p1[0][q] = ap[q];
    f4 c1[1][NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) c1[0][cb] = acc[cb];
    mma_chunk<MODE, NCB, 1>(sm, buf, p, row, grp, a1, p1, c1);
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[cb] = c1[0][cb];
}

// ============================================================ Legendre synthesis
//
// Workgroup = (order m, group of 64 rings, 32 complex columns).  Wave w owns ring tile
// 4*blk+w (16 rings) and all 64 real columns, for both parities: 8 accumulators.
// Degrees l = m + 2l' + p stream in chunks of 32 l' per parity; a staging thread loads the
// degree pairs (2g + 16i, 2g + 16i + 1), i = 0..3, of one complex column (kk = stage_column(tid),
// g = tid%8), so each load instruction reads 128 contiguous bytes per column, and stages the
// l' = g + 8i of the chunk at the chunk's K positions 4g + i: the plan stores Λ with the same
// order of K within every chunk (sht_plan.cpp: syn_kpos).


// Per-ring epilogue factors of ring rb < 64 of a synthesis block: 16 bytes in the padding of
// staging row rb (parity 0, buffer 0; the FP32 row, or the lo row of the split modes), which
// the staging and the products never touch: the Fourier row offset, then the row scale.
// Degree factor of a fused Laplacian: lap 1: e_l = -l(l+1); lap 2: 1 / e_l, 0 at l = 0.
__device__ __forceinline__ float lap_factor(int lap, int l) {
    const float e = -(float)l * (float)(l + 1);
    return lap == 1 ? e : (l == 0 ? 0.f : 1.f / e);
}

template <int MODE>
__device__ inline unsigned char *ep_slot(unsigned char *sm, int rb) {
    if constexpr (MODE == kFP32)
        return sm + (rb * kRowF + kChunk) * 4;
    else
        return sm + (kTileH + rb * kRowH + kChunk) * 2;
}

// Four workgroups per CU: at most 128 registers per lane, and the per-ring epilogue factors
// sit in the staging rows' padding, so LDS stays at 40 KiB (36 KiB in FP32).
template <int MODE, bool BATCH>
__global__ __launch_bounds__(256, MODE == kBF16x3 ? 2 : 4) void legendre_synthesis_kernel(const DevPlan *__restrict__ plans, CallArgs single,
                                                                 const CallArgs *__restrict__ calls,
                                                                 const Tile *__restrict__ tiles) {
    // the wave index is uniform per wave: keep it (and all tile ownership) scalar
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Tile T = tiles[blockIdx.x];
    if (T.m < 0) return;  // padding of the launch order
    const SHT_GLOBAL ColRef *cmap = merged_columns<BATCH>(T, single);
    const int call0 = cmap ? column_ref(cmap, 0).call : T.call;
    const CallArgs C = BATCH ? calls[call0] : single;
    const DevPlan &P = plans[C.plan];
    const int m = T.m, cg = T.cg;
    const SHT_GLOBAL float2 *coeffs = as_global(static_cast<const float2 *>(C.in));
    const long long ld_spec = C.ld_in;
    SHT_GLOBAL float4 *fbuf = as_global(static_cast<float4 *>(C.fbuf));
    const int K = C.K, unscale = C.unscale;
    const int n = P.n;
    const int j0 = as_global(P.j0)[m];
    const int J = n - j0;
    const int L = P.lmax - m;
    const int Le = (L + 1) >> 1;
    const int nch = (Le + kChunk - 1) / kChunk;
    const int Jt = (J + 15) >> 4;
    const int tile = T.blk * 4 + w;
    const bool active = tile < Jt;
    const long long off = as_global(P.lm_off)[m];
    const int kbase = cg * kColsComplex;
    const int row = lane & 15, grp = lane >> 4;
    // staging role
    const int kk = stage_column<MODE>(tid), g = tid & 7;
    bool kok = kbase + kk < K;
    const SHT_GLOBAL float2 *csrc = coeffs + (long long)(kok ? kbase + kk : 0) * ld_spec + off;
    if (BATCH && cmap) {  // the column's own call
        const ColRef r = column_ref(cmap, kk);
        const SHT_GLOBAL CallArgs *ck = call_of(calls, r.call, call0);
        kok = r.call >= 0;
        csrc = as_global(static_cast<const float2 *>(ck->in)) + (long long)(kok ? r.col : 0) * ck->ld_in + off;
    }

    __shared__ alignas(16) unsigned char sm[LegSmem<MODE>::bytes];

    f4 acc[2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[p][cb] = f4{0.f, 0.f, 0.f, 0.f};

    // Operands of one chunk: 8 degrees of one column (staging role) and this wave's Λ
    // fragments.  Two sets alternate (no register copies between chunks); the next chunk's
    // set is loaded while the current chunk's products run.  All loads are unconditional
    // (degrees past L re-read degree L - 1, absent columns column 0, idle waves their last
    // ring tile, the step after the last chunk the last chunk again; the staging zeroes what
    // is not data), so each wait covers exactly the loads it needs.
    struct Ops {
        float2 v[8];
        float af[2][8];
        V8of<MODE> ap[2][3];  // [parity][part]
    };
    const int ltile = active ? tile : Jt - 1;
    const long long so = as_global(P.syn_off)[m];
    const SHT_GLOBAL float *Af[2];
    const SHT_GLOBAL V8of<MODE> *Ap[2][3];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        if constexpr (MODE == kFP32) {
            Af[p] = as_global(P.syn_f32) + so + (long long)(p * Jt + ltile) * (nch * kChunk) * 16 + grp * 16 + row;
        } else {
            const long long fi = (long long)(p * Jt + ltile) * nch * 64 + lane;
#pragma unroll
            for (int q = 0; q < split_parts(MODE); ++q)
                Ap[p][q] = as_global(reinterpret_cast<const V8of<MODE> *>(
                               static_cast<const u16 *>(lam_part<MODE>(P.syn_hi, P.syn_mid, P.syn_lo, q)) + so)) +
                           fi;
        }
    }
    auto load = [&](Ops &S, int c) {
        c = min(c, nch - 1);
#pragma unroll
        for (int e = 0; e < 8; ++e) S.v[e] = gload(csrc + min(64 * c + 16 * (e >> 1) + 2 * g + (e & 1), L - 1));
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            if constexpr (MODE == kFP32) {
#pragma unroll
                for (int ks = 0; ks < kChunk / 4; ++ks) S.af[p][ks] = Af[p][c * kChunk * 16 + ks * 64];
            } else {
#pragma unroll
                for (int q = 0; q < split_parts(MODE); ++q) S.ap[p][q] = Ap[p][q][c * 64];
            }
        }
    };
    // stage chunk c from `cur`, load chunk c + 1 into `nxt`, then the products of chunk c
    auto step = [&](Ops &cur, Ops &nxt, int c) {
        const int buf = c & 1;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            float re[4], im[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bool ok = kok && 64 * c + 16 * i + 2 * g + p < L;
                re[i] = ok ? cur.v[2 * i + p].x : 0.f;
                im[i] = ok ? cur.v[2 * i + p].y : 0.f;
            }
            if (C.lap) {  // fused Laplacian of the input degrees (uniform)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float f = C.lap_sign * C.lap_scale * lap_factor(C.lap, m + 64 * c + 16 * i + 2 * g + p);
                    re[i] *= f;
                    im[i] *= f;
                }
            }
            stage4<MODE>(sm, buf, p, kk, g, re, im);
        }
        load(nxt, c + 1);
        __syncthreads();
        if (active) {
#pragma unroll
            for (int p = 0; p < 2; ++p)
                mma_chunk1<MODE, 4>(sm, buf, p, row, grp, cur.af[p], cur.ap[p], acc[p]);
        }
    };

    // straight-line loop body (a conditional second step would make the waits at the loop
    // head cover every outstanding load)
    // Per-ring epilogue factors of this block's 64 rings (Fourier row offset, FP16 row scale,
    // 1 / cos(lat)), loaded before the first chunk and kept in LDS until the epilogue.
    static_assert(kRowF - kChunk >= 4 && kRowH - kChunk >= 8, "16 bytes of row padding");
    long long e_row = 0;
    float e_sc = 1.f, e_ci = 1.f;
    if (w == 0) {
        const int r = min(64 * T.blk + lane, J - 1);  // ring j0 + r
        e_row = as_global(P.fboff)[j0 + r] + m;
        if constexpr (MODE == kFP16x3) e_sc = as_global(P.syn_sinv)[as_global(P.syn_soff)[m] + r];
        if (unscale) e_ci = as_global(P.coslat_inv)[j0 + r];
    }
    Ops S0, S1;
    load(S0, 0);
    if (w == 0) {
        unsigned char *slot = ep_slot<MODE>(sm, lane);
        *reinterpret_cast<long long *>(slot) = e_row;
        *reinterpret_cast<float *>(slot + 8) = e_sc * e_ci;
    }
    int c = 0;
    for (; c + 1 < nch; c += 2) {
        step(S0, S1, c);
        step(S1, S0, c + 1);
    }
    if (c < nch) step(S0, S1, c);

    if (!active) return;  // no further barriers below
    // epilogue: N = E + O, S = E - O (legendre.jl:84-91), optional 1/cos(lat)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int col = cb * 16 + row;
        const int kc = col >> 1, reim = col & 1;
        int k = kbase + kc, Kk = K;
        bool kin = k < K;
        SHT_GLOBAL float4 *fb = fbuf;
        if (BATCH && cmap) {  // the column's own call
            const ColRef r = column_ref(cmap, kc);
            const SHT_GLOBAL CallArgs *ck = call_of(calls, r.call, call0);
            kin = r.call >= 0;
            k = r.col;
            Kk = ck->K;
            fb = as_global(static_cast<float4 *>(ck->fbuf));
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 4 * grp + i;
            const int j = j0 + 16 * tile + r;
            const int rb = 16 * w + r;  // ring within the block
            const unsigned char *slot = ep_slot<MODE>(sm, rb);
            const float sc = *reinterpret_cast<const float *>(slot + 8);
            const float e = acc[0][cb][i] * sc, o = acc[1][cb][i] * sc;
            const float nv = e + o, sv = e - o;
            const float n_other = lane_xor1(nv);
            const float s_other = lane_xor1(sv);
            if (j < n && kin) {
                SHT_GLOBAL float2 *dst = reinterpret_cast<SHT_GLOBAL float2 *>(
                    fb + *reinterpret_cast<const long long *>(slot) * (long long)Kk + k);
                if (reim == 0) {
                    gstore(dst, make_float2(nv, n_other));
                } else {
                    gstore(dst + 1, make_float2(s_other, sv));
                }
            }
        }
    }
}

// ============================================================== Legendre analysis
//
// Workgroup = (order m, group of DPW = 64 TPW degrees, 32 complex columns).  Wave w owns
// parity p = w&1 and TPW consecutive degree tiles (16 rows of l' each, l = m + 2l' + p) for
// all 64 real columns, so the staged Fourier rows of a chunk serve DPW degrees.  Rings
// stream in chunks of 32; a staging thread loads 4 consecutive rings of one complex column
// (kk = stage_column(tid), rings 4(tid%8)..+3).  Every output coefficient is written exactly once (no
// atomics, deterministic).

// At least three workgroups per CU (at most 168 registers per lane).
template <int MODE, bool BATCH>
__global__ __launch_bounds__(256, MODE == kBF16x3 ? 2 : 3) void legendre_analysis_kernel(const DevPlan *__restrict__ plans, CallArgs single,
                                                                const CallArgs *__restrict__ calls,
                                                                const Tile *__restrict__ tiles) {
    constexpr int TPW = ana_tiles_per_wave(MODE);
    constexpr int DPW = ana_degs_per_wg(MODE);
    constexpr int kAnaTiles = DPW / 32;  // degree tiles per parity per workgroup
    // the wave index is uniform per wave: keep it (and all tile ownership) scalar
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Tile T = tiles[blockIdx.x];
    if (T.m < 0) return;  // padding of the launch order
    const SHT_GLOBAL ColRef *cmap = merged_columns<BATCH>(T, single);
    const int call0 = cmap ? column_ref(cmap, 0).call : T.call;
    const CallArgs C = BATCH ? calls[call0] : single;
    const DevPlan &P = plans[C.plan];
    const int m = T.m, cg = T.cg;
    const SHT_GLOBAL float4 *fbuf = as_global(static_cast<const float4 *>(C.fbuf));
    SHT_GLOBAL float2 *coeffs = as_global(static_cast<float2 *>(C.out));
    const long long ld_spec = C.ld_out;
    const int K = C.K;
    const int n = P.n;
    const int j0 = as_global(P.j0)[m];
    const int J = n - j0;
    const int L = P.lmax - m;
    const int NG = (L + DPW - 1) / DPW;
    const int nt = kAnaTiles * NG;  // degree tiles per parity in the table
    const int JC = (J + kChunk - 1) / kChunk;
    const int p = w & 1;
    const int lt0 = kAnaTiles * T.blk + TPW * (w >> 1);  // first degree tile of this wave
    const long long off = as_global(P.lm_off)[m];
    const int kbase = cg * kColsComplex;
    const int row = lane & 15, grp = lane >> 4;
    const int kk = stage_column<MODE>(tid), g = tid & 7;
    bool kok = kbase + kk < K;
    int kgl = kok ? kbase + kk : 0, sK = K;  // staging column, its Fourier scratch and row length
    const SHT_GLOBAL float4 *sfb = fbuf;
    if (BATCH && cmap) {  // the column's own call
        const ColRef r = column_ref(cmap, kk);
        const SHT_GLOBAL CallArgs *ck = call_of(calls, r.call, call0);
        kok = r.call >= 0;
        kgl = kok ? r.col : 0;
        sK = ck->K;
        sfb = as_global(static_cast<const float4 *>(ck->fbuf));
    }

    constexpr int OUTB = kColsComplex * (DPW + 1) * 8;
    constexpr int SMB = LegSmem<MODE>::bytes > OUTB ? LegSmem<MODE>::bytes : OUTB;
    __shared__ alignas(16) unsigned char sm[SMB];

    f4 acc[TPW][4];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[t][cb] = f4{0.f, 0.f, 0.f, 0.f};
    // FP16 row scales of this lane's output rows, loaded now for the epilogue
    float rs[TPW][4];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            rs[t][i] = MODE == kFP16x3 ? as_global(P.ana_sinv)[as_global(P.ana_soff)[m] + (p * nt + lt0 + t) * 16 + 4 * grp + i]
                                       : 1.f;

    // Operands of one chunk: 4 rings of one column (staging role) and this wave's Λ
    // fragments.  Two sets alternate (no register copies between chunks); the next chunk's
    // set is loaded while the current chunk's products run.  All loads are unconditional
    // (rings past n re-read ring n - 1, absent columns column 0, the step after the last
    // chunk the last chunk again; the staging zeroes what is not data), so each wait covers
    // exactly the loads it needs.  A set also carries the Fourier row offsets of the chunk
    // after its own, loaded ahead of its rows, so no Fourier load waits for the offset table.
    struct Ops {
        float4 v[4];
        float af[TPW][8];
        V8of<MODE> ap[TPW][3];  // [tile][part]
        int fo[4];  // first Fourier rows of the following chunk's rings (< 2^31 rows)
    };
    const long long ao = as_global(P.ana_off)[m];
    const SHT_GLOBAL float *Af[TPW];
    const SHT_GLOBAL V8of<MODE> *Ap[TPW][3];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int lt = lt0 + t;
        Af[t] = as_global(P.ana_f32) + ao + (long long)(p * nt + lt) * (JC * kChunk) * 16 + grp * 16 + row;
        const long long fi = (long long)(p * nt + lt) * JC * 64 + lane;
#pragma unroll
        for (int q = 0; q < split_parts(MODE); ++q)
            Ap[t][q] = as_global(reinterpret_cast<const V8of<MODE> *>(
                           static_cast<const u16 *>(lam_part<MODE>(P.ana_hi, P.ana_mid, P.ana_lo, q)) + ao)) +
                       fi;
    }
    auto load_offsets = [&](int (&fo)[4], int c) {
        c = min(c, JC - 1);
#pragma unroll
        for (int i = 0; i < 4; ++i) fo[i] = (int)as_global(P.fboff)[min(j0 + kChunk * c + 4 * g + i, n - 1)];
    };
    // chunk c into S, given the ring offsets fo of chunk c
    auto load = [&](Ops &S, int c, const int (&fo)[4]) {
        c = min(c, JC - 1);
        load_offsets(S.fo, c + 1);
#pragma unroll
        for (int i = 0; i < 4; ++i) S.v[i] = gload(sfb + (long long)(fo[i] + m) * sK + kgl);
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            if constexpr (MODE == kFP32) {
#pragma unroll
                for (int ks = 0; ks < kChunk / 4; ++ks) S.af[t][ks] = Af[t][c * kChunk * 16 + ks * 64];
            } else {
#pragma unroll
                for (int q = 0; q < split_parts(MODE); ++q) S.ap[t][q] = Ap[t][q][c * 64];
            }
        }
    };
    // stage chunk c from `cur`, load chunk c + 1 into `nxt`, then the products of chunk c
    auto step = [&](Ops &cur, Ops &nxt, int c) {
        const int buf = c & 1;
        float re[4], im[4];
        bool ok[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ok[i] = kok && j0 + kChunk * c + 4 * g + i < n;
            re[i] = ok[i] ? cur.v[i].x : 0.f;
            im[i] = ok[i] ? cur.v[i].y : 0.f;
        }
        stage4<MODE>(sm, buf, 0, kk, g, re, im);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            re[i] = ok[i] ? cur.v[i].z : 0.f;
            im[i] = ok[i] ? cur.v[i].w : 0.f;
        }
        stage4<MODE>(sm, buf, 1, kk, g, re, im);
        load(nxt, c + 1, cur.fo);
        __syncthreads();
        mma_chunk<MODE, 4, TPW>(sm, buf, p, row, grp, cur.af, cur.ap, acc);
    };

    if (JC > 0) {  // no rings for this order: the coefficients are zero
        Ops S0, S1;
        int fo0[4];
        load_offsets(fo0, 0);
        load(S0, 0, fo0);
        int c = 0;
        for (; c + 1 < JC; c += 2) {  // straight-line body, as in the synthesis
            step(S0, S1, c);
            step(S1, S0, c + 1);
        }
        if (c < JC) step(S0, S1, c);
    }
    __syncthreads();
    // stage the DPW-degree x 32-column output tile, then write l-contiguous columns
    float2 *Out = reinterpret_cast<float2 *>(sm);  // [k][l - m - DPW blk], row DPW + 1
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            const int col = cb * 16 + row;
            const int kc = col >> 1, reim = col & 1;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int lq = 16 * (TPW * (w >> 1) + t) + 4 * grp + i;  // l' - DPW / 2 blk
                float *o = reinterpret_cast<float *>(&Out[kc * (DPW + 1) + 2 * lq + p]);
                o[reim] = acc[t][cb][i] * rs[t][i];
            }
        }
    __syncthreads();
    constexpr int NIT = kColsComplex * DPW / 256;
    if (!C.accumulate && !C.lap) {  // plain transform (uniform); batch calls are plain
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = it * 256 + tid;
            const int kc = idx / DPW, o = idx % DPW;
            const int og = DPW * T.blk + o;
            int kq = kbase + kc;
            bool kin = kq < K;
            SHT_GLOBAL float2 *dst = coeffs;
            long long ld = ld_spec;
            if (BATCH && cmap) {  // the column's own call
                const ColRef r = column_ref(cmap, kc);
                const SHT_GLOBAL CallArgs *ck = call_of(calls, r.call, call0);
                kin = r.call >= 0;
                kq = r.col;
                dst = as_global(static_cast<float2 *>(ck->out));
                ld = ck->ld_out;
            }
            if (og < L && kin) gstore(dst + (long long)kq * ld + off + og, Out[kc * (DPW + 1) + o]);
        }
        return;
    }
    // accumulate: every old output of this thread is read before the first store (a store
    // between them would keep the compiler from issuing the later reads early)
    float2 old[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int idx = it * 256 + tid;
        const int kc = idx / DPW, o = idx % DPW;
        const int og = DPW * T.blk + o, kq = kbase + kc;
        old[it] = make_float2(0.f, 0.f);
        if (C.accumulate && og < L && kq < K) old[it] = gload(coeffs + (long long)kq * ld_spec + off + og);
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int idx = it * 256 + tid;
        const int kc = idx / DPW, o = idx % DPW;
        const int og = DPW * T.blk + o, kq = kbase + kc;
        if (og < L && kq < K) {
            float2 v = Out[kc * (DPW + 1) + o];
            if (C.lap) {  // fused Laplacian: sign e_l v (+ old), then the radius scaling
                const float f = C.lap_sign * lap_factor(C.lap, m + og);
                v = make_float2(f * v.x + old[it].x, f * v.y + old[it].y);
                v = make_float2(C.lap_scale * v.x, C.lap_scale * v.y);
            } else {  // pullback: add into the caller's cotangent
                v = make_float2(v.x + old[it].x, v.y + old[it].y);
            }
            gstore(coeffs + (long long)kq * ld_spec + off + og, v);
        }
    }
}

// ================================================================== entry points

template <int MODE>
static const void *legendre_entry(int stage, bool batch) {
    if (stage == kLegendreSynthesis)
        return batch ? reinterpret_cast<const void *>(&legendre_synthesis_kernel<MODE, true>)
                     : reinterpret_cast<const void *>(&legendre_synthesis_kernel<MODE, false>);
    return batch ? reinterpret_cast<const void *>(&legendre_analysis_kernel<MODE, true>)
                 : reinterpret_cast<const void *>(&legendre_analysis_kernel<MODE, false>);
}

const void *legendre_kernel(int stage, int mode, bool batch) {
    if (mode == kFP32) return legendre_entry<kFP32>(stage, batch);
    if (mode == kFP16x3) return legendre_entry<kFP16x3>(stage, batch);
    return legendre_entry<kBF16x3>(stage, batch);
}

}  // namespace sht
