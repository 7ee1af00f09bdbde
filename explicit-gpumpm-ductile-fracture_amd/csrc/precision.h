// precision.h — storage codecs of the mixed-precision particle state.
//   bf16: float with the low 16 mantissa bits dropped, round to nearest even
//   fp16: IEEE binary16 (used for F - I, whose entries stay small)
// Both are plain uint16 bit patterns in HBM; arithmetic is always FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpmx {

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    const uint32_t rne = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    const uint32_t qnan = (u >> 16) | 0x40u;
    // a select, not a branch: kernels that store many values keep one block
    return (uint16_t)(((u & 0x7fffffffu) > 0x7f800000u) ? qnan : rne);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
__host__ __device__ __forceinline__ uint16_t f2h(float f) {
    const _Float16 h = (_Float16)f;
    uint16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}
__host__ __device__ __forceinline__ float h2f(uint16_t u) {
    _Float16 h;
    __builtin_memcpy(&h, &u, 2);
    return (float)h;
}

}  // namespace mpmx
