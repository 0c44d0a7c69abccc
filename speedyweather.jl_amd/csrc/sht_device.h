// Device-side helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "sht_internal.h"

namespace sht {


typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;
typedef u16 u16x8 __attribute__((ext_vector_type(8)));

// Global-memory view of a generic pointer.  Buffers reached through structures (plan
// tables, call arguments) are generic pointers to the compiler, which then emits FLAT
// accesses that count against both the memory and the LDS counters and serialise the
// loads around every LDS access; the cast lets it emit GLOBAL accesses instead.
#define SHT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const SHT_GLOBAL T *as_global(const T *p) {
    return (const SHT_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ SHT_GLOBAL T *as_global(T *p) {
    return (SHT_GLOBAL T *)p;
}
// float2 / float4 are classes whose members take generic `this`; global accesses of them go
// through the equivalent builtin vectors.
typedef float fv2 __attribute__((ext_vector_type(2)));
typedef float fv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float2 gload(const SHT_GLOBAL float2 *p) {
    const fv2 v = *reinterpret_cast<const SHT_GLOBAL fv2 *>(p);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 gload(const SHT_GLOBAL float4 *p) {
    const fv4 v = *reinterpret_cast<const SHT_GLOBAL fv4 *>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore(SHT_GLOBAL float2 *p, float2 v) {
    *reinterpret_cast<SHT_GLOBAL fv2 *>(p) = fv2{v.x, v.y};
}
__device__ __forceinline__ void gstore(SHT_GLOBAL float4 *p, float4 v) {
    *reinterpret_cast<SHT_GLOBAL fv4 *>(p) = fv4{v.x, v.y, v.z, v.w};
}

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


__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Lane permutations within rows of 16 lanes by DPP moves (no LDS round trip, unlike
// __shfl_xor): the value of lane i ^ 1, of lane i ^ 2, and of lane 7 - (i & 7) within each
// group of 8 lanes.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_xor1(float x) { return dpp_mov<0xB1>(x); }     // quad_perm [1,0,3,2]
__device__ __forceinline__ float lane_xor2(float x) { return dpp_mov<0x4E>(x); }     // quad_perm [2,3,0,1]
__device__ __forceinline__ float lane_mirror8(float x) { return dpp_mov<0x141>(x); }  // row_half_mirror

// Exact x / d for 0 <= x < 2^16 and 1 <= d < 2^16 by a multiply-high with ceil(2^32 / d).
struct FastDiv {
    unsigned mul;
    int d;
    __device__ explicit FastDiv(int dd) : mul(dd > 1 ? 0xFFFFFFFFu / (unsigned)dd + 1u : 0u), d(dd) {}
    __device__ FastDiv(int dd, unsigned m) : mul(m), d(dd) {}  // m = fastdiv_mul(dd)
    __device__ int div(int x) const { return d == 1 ? x : (int)__umulhi((unsigned)x, mul); }
};

static inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace sht
