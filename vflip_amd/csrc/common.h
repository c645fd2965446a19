// Shared device/host helpers of libvflip (internal header).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/vflip.h"

namespace vflip {

void set_error(const char* fmt, ...);

#define VF_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::vflip::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e)); \
            return VFLIP_ERR_HIP;                                                           \
        }                                                                                   \
    } while (0)

#define VF_CHECK_LAUNCH()                                                                   \
    do {                                                                                    \
        hipError_t _e = hipGetLastError();                                                  \
        if (_e != hipSuccess) {                                                             \
            ::vflip::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
            return VFLIP_ERR_HIP;                                                           \
        }                                                                                   \
    } while (0)

#define VF_ARG(cond, msg)                                                                   \
    do {                                                                                    \
        if (!(cond)) {                                                                      \
            ::vflip::set_error("invalid argument: %s", msg);                                \
            return VFLIP_ERR_ARG;                                                           \
        }                                                                                   \
    } while (0)

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

__host__ __device__ inline size_t lin3(int i, int j, int k, int nx, int ny) {
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
}

// randhash of the reference (S/util.h:238-246): integer hash used as a
// counter-based random source.
__host__ __device__ inline unsigned randhash(unsigned seed) {
    unsigned i = (seed ^ 0xA3C59AC3u) * 2654435769u;
    i ^= (i >> 16);
    i *= 2654435769u;
    i ^= (i >> 16);
    i *= 2654435769u;
    return i;
}

// Wave64 reductions (CDNA wavefront = 64 lanes).
template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T> __device__ inline T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = v > w ? v : w;
    }
    return v;
}

// Workgroup barrier that orders LDS only: outstanding global loads stay in
// flight across it (a full __syncthreads() also waits for them).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Block reductions for NT threads (kBlock by default) in any (x, y) shape;
// result valid in the first thread.
__device__ inline int flat_tid() { return threadIdx.x + threadIdx.y * blockDim.x; }

template <typename T, int NT = kBlock> __device__ inline T block_sum(T v) {
    __shared__ T red[NT / 64];
    v = wave_sum(v);
    const int t = flat_tid(), lane = t & 63, wid = t >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T r = 0;
    if (t == 0) {
        for (int q = 0; q < NT / 64; ++q) r += red[q];
    }
    __syncthreads();
    return r;
}
template <typename T, int NT = kBlock> __device__ inline T block_max(T v) {
    __shared__ T red[NT / 64];
    v = wave_max(v);
    const int t = flat_tid(), lane = t & 63, wid = t >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T r = 0;
    if (t == 0) {
        r = red[0];
        for (int q = 1; q < NT / 64; ++q) r = r > red[q] ? r : red[q];
    }
    __syncthreads();
    return r;
}

}  // namespace vflip
