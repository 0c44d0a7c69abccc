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
