// Spectral-space operators that feed the transform (vector path of the model step).
#include "sht_device.h"

namespace sht {

// ============================================================ UV from vor, div
// spectral_gradients.jl:331-366: per packed (l, m) a 3-point stencil in l.
__global__ void uv_from_vordiv_kernel(const DevPlan *__restrict__ PP, const float2 *__restrict__ vor,
                                      const float2 *__restrict__ dv, float2 *__restrict__ U,
                                      float2 *__restrict__ V, long long ld, int K, float radius) {
    const DevPlan &P = *PP;
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

int launch_uv_from_vordiv(const DevPlan *P, int nharm, const void *vor, const void *div, void *U,
                          void *V, long long ld, int K, float radius, void *stream) {
    const long long total = (long long)nharm * K;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    uv_from_vordiv_kernel<<<blocks, 256, 0, as_stream(stream)>>>(
        P, static_cast<const float2 *>(vor), static_cast<const float2 *>(div),
        static_cast<float2 *>(U), static_cast<float2 *>(V), ld, K, radius);
    return (int)hipGetLastError();
}

}  // namespace sht
