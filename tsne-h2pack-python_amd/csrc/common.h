// common.h -- shared definitions of the gfx950 HIP implementation.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <cstddef>

#define TSNE_HIP_CHECK(expr)                                                          \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            fprintf(stderr, "tsne_h2pack: HIP error %s at %s:%d: %s\n",              \
                    hipGetErrorName(_e), __FILE__, __LINE__, #expr);                  \
            return -1;                                                                \
        }                                                                             \
    } while (0)

constexpr int kWave = 64;  // CDNA wavefront width

// Device-side interleaved coordinate record of one point in the float32 gather copy
// of the embedding: 2D uses the first two lanes of a float2, 3D a float4 (w = 0) so
// that one neighbour is a single 8- or 16-byte load.
template <int D> struct Vec;
template <> struct Vec<2> { using T = float2; };
template <> struct Vec<3> { using T = float4; };

__device__ __forceinline__ float wave_sum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// 1 / x in float32 through the hardware reciprocal (1 ulp); the kernel weights
// need ~1e-6 relative, far inside the 1e-2 H2 tolerance.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
