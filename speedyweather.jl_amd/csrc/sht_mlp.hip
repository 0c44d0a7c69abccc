// Per-column MLP of the learned tendency correction: 32 -> 128 -> 128 -> 32 with GELU,
// BF16 weights (FP32 master copy), FP32 accumulation.  Inputs are the 32 values of one
// grid column (8 levels x {u, v, T, ln ps}); outputs the 32 tendency corrections.
//
// Forward runs on MFMA (v_mfma_f32_16x16x32_bf16): activations are rounded to BF16 at each
// layer input, products accumulate in FP32, bias and GELU in FP32.  Backward differentiates
// exactly that function in FP32 arithmetic on the same BF16-rounded operands; weight
// gradients are reduced over row blocks in a fixed order (deterministic), then Adam updates
// the FP32 master weights and refreshes their BF16 copies.
//
// Parameter vector (float, kMlpParams): W1[128][32] b1[128] W2[128][128] b2[128]
// W3[32][128] b3[32], W[o][i] maps input i to output o.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>
#include <vector>

#include "../../include/sht.h"
#include "sht_device.h"

namespace sht {

constexpr int kIn = 32, kHid = 128, kOut = 32;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kOut * kHid, kMlpParams = kOffB3 + kOut;
constexpr int kRows = 64;  // rows (columns of the grid) per workgroup

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}
__device__ inline float rbf(float x) { return (float)(bf16)x; }  // round to BF16 and back

__device__ inline f4 mfma_bf16(bf16x8 a, bf16x8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// One dense layer on this wave's 16 rows: act (LDS, [kRows][din] BF16) x W^T (global BF16
// [dout][din]) + b, written as FP32 pre-activations z (LDS, [kRows][dout]).
template <int DIN, int DOUT>
__device__ inline void layer_mfma(const bf16 *act, const bf16 *__restrict__ W, const float *__restrict__ b, float *z,
                                  int wave, int lane) {
    const int row = lane & 15, grp = lane >> 4;
    constexpr int NT = DOUT / 16, KS = DIN / 32;
    f4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(act + (16 * wave + row) * DIN + 32 * ks + 8 * grp);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 w = *reinterpret_cast<const bf16x8 *>(W + (16 * t + row) * DIN + 32 * ks + 8 * grp);
            acc[t] = mfma_bf16(a, w, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 16 * wave + 4 * grp + i, o = 16 * t + row;
            z[r * DOUT + o] = acc[t][i] + b[o];
        }
}

// Forward: y = W3 g(W2 g(W1 x + b1) + b2) + b3 for rows [blk*64, blk*64+64).  Saves the
// pre-activations z1, z2 (FP32, B x 128) when given (training).
__global__ __launch_bounds__(256) void mlp_forward_kernel(const bf16 *__restrict__ Wb, const float *__restrict__ P,
                                                          const float *__restrict__ X, float *__restrict__ Y,
                                                          float *__restrict__ Z1, float *__restrict__ Z2, int B) {
    __shared__ alignas(16) bf16 a0[kRows * kIn];
    __shared__ alignas(16) bf16 a1[kRows * kHid];
    __shared__ alignas(16) float z[kRows * kHid];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * kRows;
    for (int i = tid; i < kRows * kIn; i += 256) {
        const int r = r0 + i / kIn;
        a0[i] = (bf16)(r < B ? X[(long long)r0 * kIn + i] : 0.f);
    }
    __syncthreads();
    layer_mfma<kIn, kHid>(a0, Wb + kOffW1, P + kOffB1, z, wave, lane);
    __syncthreads();
    for (int i = tid; i < kRows * kHid; i += 256) {
        const int r = r0 + i / kHid;
        if (Z1 && r < B) Z1[(long long)r0 * kHid + i] = z[i];
        a1[i] = (bf16)gelu(z[i]);
    }
    __syncthreads();
    layer_mfma<kHid, kHid>(a1, Wb + kOffW2, P + kOffB2, z, wave, lane);
    __syncthreads();
    for (int i = tid; i < kRows * kHid; i += 256) {
        const int r = r0 + i / kHid;
        if (Z2 && r < B) Z2[(long long)r0 * kHid + i] = z[i];
        a1[i] = (bf16)gelu(z[i]);
    }
    __syncthreads();
    layer_mfma<kHid, kOut>(a1, Wb + kOffW3, P + kOffB3, z, wave, lane);
    __syncthreads();
    for (int i = tid; i < kRows * kOut; i += 256) {
        const int r = r0 + i / kOut;
        if (r < B) Y[(long long)r0 * kOut + i] = z[i];
    }
}

// Backward of the mean squared error L = mean((y - t)^2) over B x 32 values for rows of one
// workgroup: writes this block's partial parameter gradient (kMlpParams floats) and its
// partial loss.  FP32 arithmetic on the BF16-rounded weights and layer inputs of forward.
__global__ __launch_bounds__(256) void mlp_backward_kernel(const float *__restrict__ P, const float *__restrict__ X,
                                                           const float *__restrict__ Y, const float *__restrict__ Tg,
                                                           const float *__restrict__ Z1, const float *__restrict__ Z2,
                                                           float *__restrict__ G, float *__restrict__ Lp, int B) {
    __shared__ float dy[kRows * kOut];
    __shared__ float h[kRows * kHid];   // layer input (BF16-rounded) of the current layer
    __shared__ float dz[kRows * kHid];  // gradient w.r.t. the current layer's pre-activation
    __shared__ float red[256];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * kRows;
    const int nr = min(kRows, B - r0);
    float *g = G + (long long)blockIdx.x * kMlpParams;
    const float scale = 2.f / ((float)B * kOut);
    float lsum = 0.f;
    for (int i = tid; i < kRows * kOut; i += 256) {
        const int r = i / kOut;
        float d = 0.f;
        if (r < nr) {
            d = Y[(long long)r0 * kOut + i] - Tg[(long long)r0 * kOut + i];
            lsum += d * d;
        }
        dy[i] = d * scale;
    }
    // layer 3 input h2 = bf16(gelu(z2))
    for (int i = tid; i < kRows * kHid; i += 256) {
        const int r = i / kHid;
        h[i] = r < nr ? rbf(gelu(Z2[(long long)r0 * kHid + i])) : 0.f;
    }
    __syncthreads();
    // dW3[o][j] = sum_r dy[r][o] h2[r][j]; db3[o] = sum_r dy[r][o]
    for (int p = tid; p < kOut * kHid; p += 256) {
        const int o = p / kHid, jj = p % kHid;
        float s = 0.f;
        for (int r = 0; r < nr; ++r) s = fmaf(dy[r * kOut + o], h[r * kHid + jj], s);
        g[kOffW3 + p] = s;
    }
    if (tid < kOut) {
        float s = 0.f;
        for (int r = 0; r < nr; ++r) s += dy[r * kOut + tid];
        g[kOffB3 + tid] = s;
    }
    // dz2[r][j] = (sum_o dy[r][o] W3b[o][j]) gelu'(z2)
    for (int i = tid; i < kRows * kHid; i += 256) {
        const int r = i / kHid, jj = i % kHid;
        float s = 0.f;
        if (r < nr) {
            for (int o = 0; o < kOut; ++o) s = fmaf(dy[r * kOut + o], rbf(P[kOffW3 + o * kHid + jj]), s);
            s *= gelu_grad(Z2[(long long)r0 * kHid + i]);
        }
        dz[i] = s;
    }
    __syncthreads();
    // layer 2 input h1 = bf16(gelu(z1))
    for (int i = tid; i < kRows * kHid; i += 256) {
        const int r = i / kHid;
        h[i] = r < nr ? rbf(gelu(Z1[(long long)r0 * kHid + i])) : 0.f;
    }
    __syncthreads();
    for (int p = tid; p < kHid * kHid; p += 256) {
        const int o = p / kHid, jj = p % kHid;
        float s = 0.f;
        for (int r = 0; r < nr; ++r) s = fmaf(dz[r * kHid + o], h[r * kHid + jj], s);
        g[kOffW2 + p] = s;
    }
    if (tid < kHid) {
        float s = 0.f;
        for (int r = 0; r < nr; ++r) s += dz[r * kHid + tid];
        g[kOffB2 + tid] = s;
    }
    __syncthreads();
    // dz1[r][j] = (sum_o dz2[r][o] W2b[o][j]) gelu'(z1), into h (h1 no longer needed)
    for (int i = tid; i < kRows * kHid; i += 256) {
        const int r = i / kHid, jj = i % kHid;
        float s = 0.f;
        if (r < nr) {
            for (int o = 0; o < kHid; ++o) s = fmaf(dz[r * kHid + o], rbf(P[kOffW2 + o * kHid + jj]), s);
            s *= gelu_grad(Z1[(long long)r0 * kHid + i]);
        }
        h[i] = s;
    }
    __syncthreads();
    // dW1[o][i] = sum_r dz1[r][o] x_b[r][i] (x rounded to BF16 as in forward)
    for (int p = tid; p < kHid * kIn; p += 256) {
        const int o = p / kIn, ii = p % kIn;
        float s = 0.f;
        for (int r = 0; r < nr; ++r) s = fmaf(h[r * kHid + o], rbf(X[(long long)(r0 + r) * kIn + ii]), s);
        g[kOffW1 + p] = s;
    }
    if (tid < kHid) {
        float s = 0.f;
        for (int r = 0; r < nr; ++r) s += h[r * kHid + tid];
        g[kOffB1 + tid] = s;
    }
    red[tid] = lsum;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) Lp[blockIdx.x] = red[0];
}

// Sums the per-block partial gradients in block order (deterministic) and the loss.
__global__ void mlp_reduce_kernel(const float *__restrict__ Gp, float *__restrict__ G, const float *__restrict__ Lp,
                                  float *__restrict__ loss, int nblk, int B) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) {
        float s = 0.f;
        for (int b = 0; b < nblk; ++b) s += Gp[(long long)b * kMlpParams + p];
        G[p] = s;
    }
    if (p == 0) {
        float s = 0.f;
        for (int b = 0; b < nblk; ++b) s += Lp[b];
        loss[0] = s / ((float)B * kOut);
    }
}

// Adam step on the FP32 master weights, refreshing the BF16 copy.
__global__ void mlp_adam_kernel(float *__restrict__ P, bf16 *__restrict__ Wb, const float *__restrict__ G,
                                float *__restrict__ Mo, float *__restrict__ Ve, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gv = G[p];
    const float m = b1 * Mo[p] + (1.f - b1) * gv;
    const float v = b2 * Ve[p] + (1.f - b2) * gv * gv;
    Mo[p] = m;
    Ve[p] = v;
    const float w = P[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[p] = w;
    Wb[p] = (bf16)w;
}

__global__ void mlp_cast_kernel(const float *__restrict__ P, bf16 *__restrict__ Wb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) Wb[p] = (bf16)P[p];
}

}  // namespace sht

using namespace sht;

struct sht_mlp {
    int device = 0;
    int max_batch = 0;
    float *P = nullptr, *M = nullptr, *V = nullptr, *G = nullptr, *Gp = nullptr, *Lp = nullptr, *loss = nullptr;
    float *Y = nullptr, *Z1 = nullptr, *Z2 = nullptr;
    bf16 *Wb = nullptr;
    long long step = 0;
};

namespace {
int mlp_fail(int code, const std::string &m) {
    sht::set_last_error(m);
    return code;
}
struct Dev {
    int prev = -1;
    bool ch = false;
    explicit Dev(int d) {
        if (hipGetDevice(&prev) == hipSuccess && prev != d) ch = hipSetDevice(d) == hipSuccess;
    }
    ~Dev() {
        if (ch) (void)hipSetDevice(prev);
    }
};
void mlp_free(sht_mlp *m) {
    if (!m) return;
    for (void *p : {(void *)m->P, (void *)m->M, (void *)m->V, (void *)m->G, (void *)m->Gp, (void *)m->Lp,
                    (void *)m->loss, (void *)m->Y, (void *)m->Z1, (void *)m->Z2, (void *)m->Wb})
        if (p) (void)hipFree(p);
    delete m;
}
hipStream_t st(void *s) { return reinterpret_cast<hipStream_t>(s); }
}  // namespace

extern "C" {

int sht_mlp_nparams(void) { return kMlpParams; }

int sht_mlp_create(int32_t device, int32_t max_batch, sht_mlp **out) {
    if (!out || max_batch < 1) return mlp_fail(SHT_ERR_ARG, "bad arguments");
    *out = nullptr;
    sht_mlp *m = new (std::nothrow) sht_mlp();
    if (!m) return mlp_fail(SHT_ERR_ALLOC, "host allocation failed");
    m->device = device;
    m->max_batch = max_batch;
    Dev g(device);
    const int nblk = (max_batch + kRows - 1) / kRows;
    const size_t rows = (size_t)nblk * kRows;
    struct A {
        void **p;
        size_t bytes;
    } allocs[] = {{(void **)&m->P, kMlpParams * 4},
                  {(void **)&m->M, kMlpParams * 4},
                  {(void **)&m->V, kMlpParams * 4},
                  {(void **)&m->G, kMlpParams * 4},
                  {(void **)&m->Gp, (size_t)nblk * kMlpParams * 4},
                  {(void **)&m->Lp, (size_t)nblk * 4},
                  {(void **)&m->loss, 4},
                  {(void **)&m->Y, rows * kOut * 4},
                  {(void **)&m->Z1, rows * kHid * 4},
                  {(void **)&m->Z2, rows * kHid * 4},
                  {(void **)&m->Wb, kMlpParams * 2}};
    for (auto &a : allocs) {
        if (hipMalloc(a.p, a.bytes) != hipSuccess) {
            mlp_free(m);
            return mlp_fail(SHT_ERR_ALLOC, "hipMalloc failed");
        }
    }
    if (hipMemset(m->M, 0, kMlpParams * 4) != hipSuccess || hipMemset(m->V, 0, kMlpParams * 4) != hipSuccess) {
        mlp_free(m);
        return mlp_fail(SHT_ERR_HIP, "hipMemset failed");
    }
    *out = m;
    return SHT_OK;
}

int sht_mlp_set_params(sht_mlp *m, const float *host_params) {
    if (!m || !host_params) return mlp_fail(SHT_ERR_ARG, "bad arguments");
    Dev g(m->device);
    if (hipMemcpy(m->P, host_params, kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess)
        return mlp_fail(SHT_ERR_HIP, "hipMemcpy failed");
    mlp_cast_kernel<<<(kMlpParams + 255) / 256, 256>>>(m->P, m->Wb);
    (void)hipMemset(m->M, 0, kMlpParams * 4);
    (void)hipMemset(m->V, 0, kMlpParams * 4);
    m->step = 0;
    return hipDeviceSynchronize() == hipSuccess ? SHT_OK : mlp_fail(SHT_ERR_HIP, "kernel failed");
}

int sht_mlp_get_params(sht_mlp *m, float *host_params) {
    if (!m || !host_params) return mlp_fail(SHT_ERR_ARG, "bad arguments");
    Dev g(m->device);
    return hipMemcpy(host_params, m->P, kMlpParams * 4, hipMemcpyDeviceToHost) == hipSuccess
               ? SHT_OK
               : mlp_fail(SHT_ERR_HIP, "hipMemcpy failed");
}

int sht_mlp_forward(sht_mlp *m, const float *x, float *y, int32_t B, void *stream) {
    if (!m || !x || !y || B < 1) return mlp_fail(SHT_ERR_ARG, "bad arguments");
    Dev g(m->device);
    const int nblk = (B + kRows - 1) / kRows;
    mlp_forward_kernel<<<nblk, 256, 0, st(stream)>>>(m->Wb, m->P, x, y, nullptr, nullptr, B);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHT_OK : mlp_fail(SHT_ERR_HIP, hipGetErrorString(e));
}

static int mlp_grad(sht_mlp *m, const float *x, const float *t, int B, void *stream) {
    const int nblk = (B + kRows - 1) / kRows;
    mlp_forward_kernel<<<nblk, 256, 0, st(stream)>>>(m->Wb, m->P, x, m->Y, m->Z1, m->Z2, B);
    mlp_backward_kernel<<<nblk, 256, 0, st(stream)>>>(m->P, x, m->Y, t, m->Z1, m->Z2, m->Gp, m->Lp, B);
    mlp_reduce_kernel<<<(kMlpParams + 255) / 256, 256, 0, st(stream)>>>(m->Gp, m->G, m->Lp, m->loss, nblk, B);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHT_OK : mlp_fail(SHT_ERR_HIP, hipGetErrorString(e));
}

int sht_mlp_gradient(sht_mlp *m, const float *x, const float *t, int32_t B, float *grad, float *loss, void *stream) {
    if (!m || !x || !t || !grad || !loss || B < 1 || B > m->max_batch) return mlp_fail(SHT_ERR_ARG, "bad arguments");
    Dev g(m->device);
    int rc = mlp_grad(m, x, t, B, stream);
    if (rc) return rc;
    if (hipMemcpyAsync(grad, m->G, kMlpParams * 4, hipMemcpyDeviceToDevice, st(stream)) != hipSuccess ||
        hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToDevice, st(stream)) != hipSuccess)
        return mlp_fail(SHT_ERR_HIP, "hipMemcpyAsync failed");
    return SHT_OK;
}

int sht_mlp_train_step(sht_mlp *m, const float *x, const float *t, int32_t B, float lr, float *loss, void *stream) {
    if (!m || !x || !t || B < 1 || B > m->max_batch) return mlp_fail(SHT_ERR_ARG, "bad arguments");
    Dev g(m->device);
    int rc = mlp_grad(m, x, t, B, stream);
    if (rc) return rc;
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
    mlp_adam_kernel<<<(kMlpParams + 255) / 256, 256, 0, st(stream)>>>(m->P, m->Wb, m->G, m->M, m->V, lr, b1, b2,
                                                                      1e-8f, c1, c2);
    if (loss && hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToDevice, st(stream)) != hipSuccess)
        return mlp_fail(SHT_ERR_HIP, "hipMemcpyAsync failed");
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? SHT_OK : mlp_fail(SHT_ERR_HIP, hipGetErrorString(e));
}

int sht_mlp_destroy(sht_mlp *m) {
    if (!m) return SHT_OK;
    Dev g(m->device);
    (void)hipDeviceSynchronize();
    mlp_free(m);
    return SHT_OK;
}

}  // extern "C"
