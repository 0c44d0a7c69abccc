// Spectral-space operators that feed the transform (vector path of the model step).
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
