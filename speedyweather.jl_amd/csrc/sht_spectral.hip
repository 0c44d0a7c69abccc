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
// C (M x K, column stride ldc) = A (M x N, column-major, column stride M) B (N x K, column
// stride ldb), then optionally C[r, k] *= row_scale[r].  The dense transform's products
// have few columns (K layers) and a large A that must be read once, so this is a split-K
// skinny GEMM bound by reading A: workgroup (row block, split, column group) reads rows
// 1024 b .. + 1023 (thread t: rows t + 256 i, i < 4, so each load instruction reads 1 KiB
// of a column) of the columns of its split, with the split's rows of B (KB columns) staged
// in LDS and read as broadcasts; partial sums go to W[s][k][r], which a second kernel adds
// up in a fixed order (no atomics, deterministic).

constexpr int kDenseRows = 1024;  // rows per workgroup (4 per thread)
constexpr int kDenseMaxNS = 256;  // columns of A per split at most (LDS for B)

// U columns of A per register set; two sets alternate, so the next set's loads are in
// flight while the current set's products run (loads past the split re-read its last
// column and are not used).
template <int KB>
__global__ __launch_bounds__(256) void dense_partial_kernel(const float *__restrict__ A_, long long M, long long N,
                                                            const float *__restrict__ B_, long long ldb, int K,
                                                            int ns, float *__restrict__ W_) {
    constexpr int U = KB == 8 ? 8 : 4;
    __shared__ float Bs[kDenseMaxNS][KB];
    const int tid = threadIdx.x;
    const long long r0 = (long long)blockIdx.x * kDenseRows + tid;
    const int sp = blockIdx.y, k0 = blockIdx.z * KB;
    const long long n0 = (long long)sp * ns;
    const int nn = (int)min((long long)ns, N - n0);  // >= 1: the host makes every split non-empty
    const SHT_GLOBAL float *B = as_global(B_);
    for (int i = tid; i < nn * KB; i += 256) {
        const int q = i / KB, k = i - q * KB;
        Bs[q][k] = k0 + k < K ? B[(n0 + q) + (long long)(k0 + k) * ldb] : 0.f;
    }
    int rr[4];  // rows past M re-read row M - 1 (unconditional loads); their sums are dropped
#pragma unroll
    for (int i = 0; i < 4; ++i) rr[i] = (int)min(r0 + 256 * i, M - 1);
    float acc[4][KB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int k = 0; k < KB; ++k) acc[i][k] = 0.f;
    const SHT_GLOBAL float *A = as_global(A_) + n0 * M;
    auto fetch = [&](float (&a)[U][4], int q) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const SHT_GLOBAL float *col = A + (long long)min(q + u, nn - 1) * M;
#pragma unroll
            for (int i = 0; i < 4; ++i) a[u][i] = col[rr[i]];
        }
    };
    auto products = [&](const float (&a)[U][4], int q) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (q + u >= nn) break;  // uniform
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const float b = Bs[q + u][k];
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[i][k] = fmaf(a[u][i], b, acc[i][k]);
            }
        }
    };
    float a0[U][4], a1[U][4];
    fetch(a0, 0);
    __syncthreads();  // B staged
    int q = 0;
    for (; q + U < nn; q += 2 * U) {  // straight-line body: each set's wait covers its own loads
        fetch(a1, q + U);
        products(a0, q);
        fetch(a0, q + 2 * U);
        products(a1, q + U);
    }
    if (q < nn) products(a0, q);
    SHT_GLOBAL float *W = as_global(W_);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const long long r = r0 + 256 * i;
        if (r >= M) continue;
#pragma unroll
        for (int k = 0; k < KB; ++k)
            if (k0 + k < K) W[((long long)sp * K + k0 + k) * M + r] = acc[i][k];
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
    d.kb = K <= 8 ? 8 : 16;
    d.rb = (int)((M + kDenseRows - 1) / kDenseRows);
    d.cg = (K + d.kb - 1) / d.kb;
    long long s = (512 + (long long)d.rb * d.cg - 1) / ((long long)d.rb * d.cg);  // about 512 workgroups
    s = std::min(s, (N + 31) / 32);                                               // at least 32 columns each
    s = std::max(s, (N + kDenseMaxNS - 1) / kDenseMaxNS);                         // at most kDenseMaxNS
    s = std::max(s, 1LL);
    d.ns = (int)((N + s - 1) / s);
    d.s = (int)((N + d.ns - 1) / d.ns);  // every split non-empty
    return d;
}

int launch_dense_gemm(const float *A, long long M, long long N, const float *B, long long ldb, int K, float *C,
                      long long ldc, const float *row_scale, float *work, void *stream) {
    if (M < 1 || N < 1 || K < 1 || K > 65535 || ldb < N || ldc < M) return (int)hipErrorInvalidValue;
    const DenseSplit d = dense_split(M, N, K);
    if (d.s > 65535 || d.cg > 65535) return (int)hipErrorInvalidValue;
    const hipStream_t st = as_stream(stream);
    const dim3 grid(d.rb, d.s, d.cg);
    if (d.kb == 8)
        dense_partial_kernel<8><<<grid, 256, 0, st>>>(A, M, N, B, ldb, K, d.ns, work);
    else
        dense_partial_kernel<16><<<grid, 256, 0, st>>>(A, M, N, B, ldb, K, d.ns, work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
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
