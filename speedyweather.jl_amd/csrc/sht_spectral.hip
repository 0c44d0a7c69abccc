// Spectral-space operators that feed the transform (vector path of the model step), and
// the dense transform's GEMM.
#include <algorithm>

#include "sht_device.h"

namespace sht {

// ============================================================ UV from vor, div
// spectral_gradients.jl:331-366: per packed (l, m) a 3-point stencil in l.
__global__ void uv_from_vordiv_kernel(const DevPlan *__restrict__ PP, const float2 *__restrict__ vor_,
                                      const float2 *__restrict__ dv_, float2 *__restrict__ U_,
                                      float2 *__restrict__ V_, long long ld, float radius) {
    const DevPlan &P = *PP;
    const int lm = blockIdx.x * blockDim.x + threadIdx.x;
    if (lm >= P.nharm) return;
    const long long base = (long long)blockIdx.y * ld;  // column k = blockIdx.y
    const SHT_GLOBAL float2 *vor = as_global(vor_) + base;
    const SHT_GLOBAL float2 *dv = as_global(dv_) + base;
    const int l = as_global(P.l_of)[lm];
    const float x = as_global(P.uv_x)[lm], u1 = as_global(P.uv_1)[lm], u2 = as_global(P.uv_2)[lm];
    const float2 z = gload(vor + lm), d = gload(dv + lm);
    float2 du = make_float2(0.f, 0.f), dvv = make_float2(0.f, 0.f);
    // i x D and i x ζ
    float2 uu, vv;
    if (lm == 0) {
        float2 zp = make_float2(0.f, 0.f), dp = make_float2(0.f, 0.f);
        if (P.lmax > 1) {
            zp = gload(vor + 1);
            dp = gload(dv + 1);
        }
        uu = make_float2(-x * d.y + u2 * zp.x, x * d.x + u2 * zp.y);
        vv = make_float2(-x * z.y - u2 * dp.x, x * z.x - u2 * dp.y);
    } else if (l == P.lmax - 1) {
        const float2 zm = gload(vor + lm - 1), dm = gload(dv + lm - 1);
        uu = make_float2(-u1 * zm.x, -u1 * zm.y);
        vv = make_float2(u1 * dm.x, u1 * dm.y);
    } else {
        const float2 zm = gload(vor + lm - 1), dm = gload(dv + lm - 1);
        if (l < P.lmax - 2) {
            du = gload(vor + lm + 1);
            dvv = gload(dv + lm + 1);
        }
        // u2 is 0 on the second-to-last row, so du, dvv only enter below it
        uu = make_float2(-x * d.y + u2 * du.x - u1 * zm.x, x * d.x + u2 * du.y - u1 * zm.y);
        vv = make_float2(-x * z.y + u1 * dm.x - u2 * dvv.x, x * z.x + u1 * dm.y - u2 * dvv.y);
    }
    gstore(as_global(U_) + base + lm, make_float2(uu.x * radius, uu.y * radius));
    gstore(as_global(V_) + base + lm, make_float2(vv.x * radius, vv.y * radius));
}

// ======================================================= spectral gradient operators
// Per packed (l, m) and column, stencils in l with the reference's factor arrays
// (gradient_arrays.jl:4-21,27-76,80-100,150-170), in float:
//   divergence (spectral_gradients.jl:105-125): i m_x u - gyvd1 v[lm-1] + gyvd2 v[lm+1]
//   curl (:30-41, the divergence kernel on (v, u) with the signs of b, c swapped)
//   Laplacian (:385-420): a * (-l(l+1)) or a / (-l(l+1)) (0 at l = 0)
//   gradient (:483-527): dpdx = i m p, dpdy = gy1 p[lm-1] + gy2 p[lm+1]
// Factors crossing an order boundary (l = m below, l = lmax-1 above) are zero by construction.
// The last packed element of the divergence / curl is 0; flipsign, add and the radius
// scaling follow the reference's order (scaling after adding).
__global__ void spectral_op_kernel(const DevPlan *__restrict__ PP, int op, const float2 *__restrict__ a_,
                                   const float2 *__restrict__ b_, float2 *__restrict__ o1_, float2 *__restrict__ o2_,
                                   long long ld, int flags, float rscale) {
    const DevPlan &P = *PP;
    const int lm = blockIdx.x * blockDim.x + threadIdx.x;
    const int last = P.nharm - 1;
    if (lm > last) return;
    const long long base = (long long)blockIdx.y * ld;
    const SHT_GLOBAL float2 *a = as_global(a_) + base;
    SHT_GLOBAL float2 *o1 = as_global(o1_) + base;
    const bool flip = flags & kOpFlip, add = flags & kOpAdd;
    float2 val;
    if (op == kOpDivergence || op == kOpCurl) {
        if (lm == last) {
            val = make_float2(0.f, 0.f);
        } else {
            const SHT_GLOBAL float2 *b = as_global(b_) + base;
            const SHT_GLOBAL float2 *X = op == kOpCurl ? b : a;  // zonal operand
            const SHT_GLOBAL float2 *Y = op == kOpCurl ? a : b;  // meridional operand
            const float4 gy = gload(as_global(P.gy) + lm);       // (gyvd1, gyvd2, gy1, gy2)
            const float mx = gload(as_global(P.gx) + lm).x;
            const float2 z = gload(X + lm);
            const float2 ym = lm > 0 ? gload(Y + lm - 1) : make_float2(0.f, 0.f);
            const float2 yp = gload(Y + lm + 1);
            const float sg = op == kOpCurl ? 1.f : -1.f;
            val = make_float2(-mx * z.y + sg * (gy.x * ym.x - gy.y * yp.x), mx * z.x + sg * (gy.x * ym.y - gy.y * yp.y));
            if (flip) val = make_float2(-val.x, -val.y);
            if (add) {
                const float2 o = gload(o1 + lm);
                val = make_float2(o.x + val.x, o.y + val.y);
            }
        }
    } else if (op == kOpLaplace) {
        const int l = as_global(P.l_of)[lm];
        float e = -(float)(l * (l + 1));
        if (flags & kOpInverse) e = l == 0 ? 0.f : 1.f / e;
        const float2 z = gload(a + lm);
        val = make_float2(z.x * e, z.y * e);
        if (flip) val = make_float2(-val.x, -val.y);
        if (add) {
            const float2 o = gload(o1 + lm);
            val = make_float2(o.x + val.x, o.y + val.y);
        }
    } else {  // gradient: o1 = dpdx, o2 = dpdy
        const float m = gload(as_global(P.gx) + lm).y;
        const float4 gy = gload(as_global(P.gy) + lm);
        const float2 z = gload(a + lm);
        const float2 pm = lm > 0 ? gload(a + lm - 1) : make_float2(0.f, 0.f);
        const float2 pp = lm < last ? gload(a + lm + 1) : make_float2(0.f, 0.f);
        gstore(o1 + lm, make_float2(-m * z.y * rscale, m * z.x * rscale));
        val = make_float2(gy.z * pm.x + gy.w * pp.x, gy.z * pm.y + gy.w * pp.y);
        gstore(as_global(o2_) + base + lm, make_float2(val.x * rscale, val.y * rscale));
        return;
    }
    gstore(o1 + lm, make_float2(val.x * rscale, val.y * rscale));
}

int launch_spectral_op(const DevPlan *P, int nharm, int op, const void *a, const void *b, void *o1, void *o2,
                       long long ld, int K, int flags, float rscale, void *stream) {
    if (nharm < 1 || K < 1 || K > 65535) return (int)hipErrorInvalidValue;
    const dim3 grid((nharm + 255) / 256, K);
    spectral_op_kernel<<<grid, 256, 0, as_stream(stream)>>>(P, op, static_cast<const float2 *>(a),
                                                            static_cast<const float2 *>(b), static_cast<float2 *>(o1),
                                                            static_cast<float2 *>(o2), ld, flags, rscale);
    return (int)hipGetLastError();
}

// =========================================================== per-point scaling
__global__ void scale_points_kernel(const float *__restrict__ scale, float *__restrict__ field, long long npoints,
                                    long long ld) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npoints) return;
    SHT_GLOBAL float *f = as_global(field) + (long long)blockIdx.y * ld + p;
    *f = *f * as_global(scale)[p];
}

int launch_scale_points(const float *scale, void *field, long long npoints, long long ld, int K, void *stream) {
    if (npoints < 1 || K < 1 || K > 65535) return (int)hipErrorInvalidValue;
    const dim3 grid((unsigned)((npoints + 255) / 256), K);
    scale_points_kernel<<<grid, 256, 0, as_stream(stream)>>>(scale, static_cast<float *>(field), npoints, ld);
    return (int)hipGetLastError();
}

// ============================================================ dense transform GEMM
// C (M x K, column stride ldc) = A (M x N) B (N x K, column stride ldb), then optionally
// C[r, k] *= row_scale[r].  The dense transform's products have few columns (K layers) and
// a large A that must be read once, so the GEMM runs on v_mfma_f32_16x16x4_f32 with A in
// fragment order (dense_fragment_index: each wave's load of four K-steps is 1 KiB
// contiguous) and is bound by reading A for K up to 32.  Workgroup = four waves, each one
// 16-row tile, times a split of the contraction; the split's rows of B (all K columns, up
// to 64 per column group) are staged in LDS 64 at a time (all of a thread's loads in flight
// before its stores) and shared by the four row tiles, and each wave's A loads for the next
// 64 rows are issued before the current ones' products.  With one split the result is stored directly; otherwise partial sums go to
// W[s][k][r] and a second kernel adds them in a fixed order (no atomics, deterministic).

constexpr int kDenseChunk = 4;  // fragment groups (16 contraction values each) staged at once
constexpr int kDenseCols = 64;  // columns of B per column group (four column tiles)

template <int NCT>
__global__ __launch_bounds__(256) void dense_mfma_kernel(const float4 *__restrict__ Af_, long long RT, long long S4,
                                                         long long M, long long N, const float *__restrict__ B_,
                                                         long long ldb, int K, int s4_split, float *__restrict__ out_,
                                                         long long ldo, int direct,
                                                         const float *__restrict__ row_scale) {
    constexpr int BST = 16 * NCT + 16;  // LDS row stride: the four lane groups' rows on distinct banks
    __shared__ float Bs[kDenseChunk * 16][BST];
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int row = lane & 15, kg = lane >> 4;
    const long long rt = (long long)blockIdx.x * 4 + w;  // idle waves (rt >= RT) still stage and synchronise
    const int sp = blockIdx.y;
    const int k0 = blockIdx.z * kDenseCols, Kg = min(kDenseCols, K - k0);
    const long long s4a = (long long)sp * s4_split, s4b = min(S4, s4a + s4_split);
    const SHT_GLOBAL float *B = as_global(B_) + (long long)k0 * ldb;
    const SHT_GLOBAL fv4 *Aw = reinterpret_cast<const SHT_GLOBAL fv4 *>(as_global(Af_)) +
                               (min(rt, RT - 1) * S4) * 64 + lane;
    f4 acc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) acc[ct] = f4{0.f, 0.f, 0.f, 0.f};
    // B staging: thread t takes rows q = (t + 256 u) % 128 of column j = (t + 256 u) / 128,
    // all loads of a chunk in flight before its LDS stores
    constexpr int CR = kDenseChunk * 16;                // rows staged per chunk
    constexpr int PER = CR * 16 * NCT / 256;            // B values per thread and chunk
    // this wave's A fragments: the next chunk's loads are issued before the current
    // chunk's products (two register sets)
    fv4 a0[kDenseChunk], a1[kDenseChunk];
    auto load_a = [&](fv4 (&a)[kDenseChunk], long long c0) {
        const long long nch = min((long long)kDenseChunk, s4b - c0);
#pragma unroll
        for (int c = 0; c < kDenseChunk; ++c) a[c] = Aw[(c0 + min((long long)c, nch - 1)) * 64];
    };
    auto chunk = [&](const fv4 (&a)[kDenseChunk], long long c0) {
        const int nch = (int)min((long long)kDenseChunk, s4b - c0);
        float bv[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = tid + 256 * u, q = i % CR, j = i / CR;
            const long long n = 16 * c0 + q;
            bv[u] = (q < 16 * nch && n < N && j < Kg) ? B[n + (long long)j * ldb] : 0.f;
        }
        __syncthreads();  // the previous chunk's B is no longer read
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = tid + 256 * u;
            Bs[i % CR][i / CR] = bv[u];
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kDenseChunk; ++c) {
            if (c >= nch) break;  // uniform
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int q = 16 * c + 4 * v + kg;
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) acc[ct] = mfma_f32(a[c][v], Bs[q][ct * 16 + row], acc[ct]);
            }
        }
    };
    long long c0 = s4a;
    if (c0 < s4b) load_a(a0, c0);
    for (; c0 < s4b; c0 += 2 * kDenseChunk) {  // uniform; straight-line body, each set waits for its own loads
        if (c0 + kDenseChunk < s4b) load_a(a1, c0 + kDenseChunk);
        chunk(a0, c0);
        if (c0 + kDenseChunk >= s4b) break;
        if (c0 + 2 * kDenseChunk < s4b) load_a(a0, c0 + 2 * kDenseChunk);
        chunk(a1, c0 + kDenseChunk);
    }
    if (rt >= RT) return;
    SHT_GLOBAL float *out = as_global(out_);
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
        const int k = ct * 16 + row;
        if (k >= Kg) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long r = rt * 16 + 4 * kg + i;
            if (r >= M) continue;
            if (direct)
                out[r + (long long)(k0 + k) * ldo] = acc[ct][i] * (row_scale ? as_global(row_scale)[r] : 1.f);
            else
                out[((long long)sp * K + k0 + k) * M + r] = acc[ct][i];
        }
    }
}

// Sum of the S partials of 32 rows of one column k: 8 thread groups each add every 8th
// split (independent loads in flight), then group 0 adds the 8 group sums in order.
constexpr int kReduceRows = 32, kReduceGroups = 8;
__global__ __launch_bounds__(256) void dense_reduce_kernel(const float *__restrict__ W_, long long M, int K, int S,
                                                           const float *__restrict__ row_scale,
                                                           float *__restrict__ C_, long long ldc) {
    __shared__ float part[kReduceGroups][kReduceRows];
    const int row = threadIdx.x % kReduceRows, grp = threadIdx.x / kReduceRows;
    const long long r = (long long)blockIdx.x * kReduceRows + row;
    const int k = blockIdx.y;
    const long long rc = r < M ? r : M - 1;  // rows past M read row M - 1 and are not stored
    const SHT_GLOBAL float *W = as_global(W_) + (long long)k * M + rc;
    const long long step = (long long)K * M;
    float sum = 0.f;
#pragma unroll 4
    for (int sp = grp; sp < S; sp += kReduceGroups) sum += W[(long long)sp * step];
    part[grp][row] = sum;
    __syncthreads();
    if (grp != 0 || r >= M) return;
    float t = part[0][row];
#pragma unroll
    for (int g = 1; g < kReduceGroups; ++g) t += part[g][row];
    if (row_scale) t *= as_global(row_scale)[r];
    as_global(C_)[r + (long long)k * ldc] = t;
}

DenseSplit dense_split(long long M, long long N, int K) {
    DenseSplit d;
    d.rt = (M + 15) / 16;
    d.s4 = (N + 15) / 16;
    d.cg = (K + kDenseCols - 1) / kDenseCols;
    const long long wgs = (d.rt + 3) / 4 * d.cg;
    long long s = (2048 + wgs - 1) / wgs;  // about 2048 workgroups
    s = std::max(1LL, std::min(s, d.s4));
    d.s4_split = (int)((d.s4 + s - 1) / s);
    d.s = (int)((d.s4 + d.s4_split - 1) / d.s4_split);  // every split non-empty
    return d;
}

long long dense_fragment_index(long long r, long long n, long long S4) {
    const long long rt = r / 16, s4 = n / 16, q = n % 16;
    const long long lane = (r % 16) + 16 * (q % 4), v = q / 4;
    return ((rt * S4 + s4) * 64 + lane) * 4 + v;
}

int launch_dense_gemm(const float *Af, long long M, long long N, const float *B, long long ldb, int K, float *C,
                      long long ldc, const float *row_scale, float *work, void *stream) {
    if (M < 1 || N < 1 || K < 1 || K > 65535 || ldb < N || ldc < M) return (int)hipErrorInvalidValue;
    const DenseSplit d = dense_split(M, N, K);
    if ((d.rt + 3) / 4 > 0x7fffffffLL || d.s > 65535 || d.cg > 65535) return (int)hipErrorInvalidValue;
    const hipStream_t st = as_stream(stream);
    const dim3 grid((unsigned)((d.rt + 3) / 4), d.s, d.cg);
    const bool direct = d.s == 1;
    float *out = direct ? C : work;
    const long long ldo = direct ? ldc : M;
    const float4 *A4 = reinterpret_cast<const float4 *>(Af);
    const int nct = std::min(4, (std::min(K, kDenseCols) + 15) / 16);
    switch (nct) {
        case 1: dense_mfma_kernel<1><<<grid, 256, 0, st>>>(A4, d.rt, d.s4, M, N, B, ldb, K, d.s4_split, out, ldo, direct, row_scale); break;
        case 2: dense_mfma_kernel<2><<<grid, 256, 0, st>>>(A4, d.rt, d.s4, M, N, B, ldb, K, d.s4_split, out, ldo, direct, row_scale); break;
        case 3: dense_mfma_kernel<3><<<grid, 256, 0, st>>>(A4, d.rt, d.s4, M, N, B, ldb, K, d.s4_split, out, ldo, direct, row_scale); break;
        default: dense_mfma_kernel<4><<<grid, 256, 0, st>>>(A4, d.rt, d.s4, M, N, B, ldb, K, d.s4_split, out, ldo, direct, row_scale); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || direct) return (int)e;
    dense_reduce_kernel<<<dim3((unsigned)((M + kReduceRows - 1) / kReduceRows), K), 256, 0, st>>>(work, M, K, d.s,
                                                                                                row_scale, C, ldc);
    return (int)hipGetLastError();
}

int launch_uv_from_vordiv(const DevPlan *P, int nharm, const void *vor, const void *div, void *U,
                          void *V, long long ld, int K, float radius, void *stream) {
    if (nharm < 1 || K < 1 || K > 65535) return (int)hipErrorInvalidValue;
    const dim3 grid((nharm + 255) / 256, K);
    uv_from_vordiv_kernel<<<grid, 256, 0, as_stream(stream)>>>(
        P, static_cast<const float2 *>(vor), static_cast<const float2 *>(div),
        static_cast<float2 *>(U), static_cast<float2 *>(V), ld, radius);
    return (int)hipGetLastError();
}

}  // namespace sht
