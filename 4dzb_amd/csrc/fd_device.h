// Device-side arithmetic of the visibility path, in the reference's
// expression order.  Compiled with -ffp-contract=off, IEEE division and
// denormals preserved, so every float result is the correctly rounded value
// of the same expression the reference evaluates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fd_types.h"

namespace fdk {

// Order-preserving float <-> uint32 encoding (-0 < +0); used for min/max
// reductions with integer atomics.  NaN must be filtered by the caller.
__device__ __forceinline__ uint32_t fenc(float x) {
    uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t e) {
    uint32_t u = (e & 0x80000000u) ? (e & 0x7fffffffu) : ~e;
    return __uint_as_float(u);
}

// min/max that ignore NaN and order -0 below +0.
__device__ __forceinline__ float fmin_(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return fdec(min(fenc(a), fenc(b)));
}
__device__ __forceinline__ float fmax_(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return fdec(max(fenc(a), fenc(b)));
}

// float -> int32 toward zero, saturating, NaN -> 0.
__device__ __forceinline__ int cvt_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int)0x80000000;
    return (int)x;
}
// float -> uint16 toward zero, saturating, NaN -> 0.
__device__ __forceinline__ int cvt_sat_u16(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 65535.0f) return 65535;
    return (int)x;
}

// Light-plane projection, FDc/FDprojection.inl:6-19.
__device__ __forceinline__ float getZ(const float* m, float x, float y, float z) {
    return m[3] * x + m[7] * y + m[11] * z + m[15] * 1.0f;
}
__device__ __forceinline__ void makeRect(const float* m, float x, float y, float z, float& rx, float& ry, float& rz) {
    rz = m[3] * x + m[7] * y + m[11] * z + m[15] * 1.0f;
    rx = m[0] * x + m[4] * y + m[8] * z + m[12] * 1.0f;
    rx = rx / rz * 0.5f + 0.5f;
    ry = m[1] * x + m[5] * y + m[9] * z + m[13] * 1.0f;
    ry = ry / rz * 0.5f + 0.5f;
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

// Wave-wide (64-lane) reductions.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// NaN-ignoring float extrema over the wave (fmaxf / fminf)
__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// rank of this lane among the set lanes of `m` below it
__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

}  // namespace fdk
