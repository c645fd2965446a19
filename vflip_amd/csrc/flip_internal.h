// Particle workspace and helpers shared by the FLIP transfer (flip.hip) and
// the nested FLIP coupling (nested_flip.hip). Not part of the C ABI.
#pragma once

#include <climits>

#include "common.h"

struct vflip_particles {
    int nx = 0, ny = 0, nz = 0;
    float h = 0;
    size_t ncell = 0;
    hipStream_t stream = nullptr;
    int* slot = nullptr;     // per particle: rank inside its cell's atomic slot range
    int* tmp = nullptr;      // per particle: merge buffer of the long-segment sort
    int* big = nullptr;      // [0]: count, [1..]: cells whose segment needs the long sort
    int* partial = nullptr;  // scan partials
    int* reo = nullptr;      // reorder: per-cell survivor counts/offsets [ncell+1], masks [ncell]
    size_t cap = 0, big_cap = 0, npartial = 0;
    int* pflag = nullptr;     // per-particle flags/offsets of a selection [pflag_cap + 1]
    int* ppartial = nullptr;  // scan partials of pflag
    size_t pflag_cap = 0;
};

namespace vflip {

// Truncating float->int of Vec3i(Vec3f) (S/vec.h:37-39); NaN and values
// outside int range give INT_MIN, the x86-64 conversion result.
__device__ inline int trunc_i32(float t) {
    if (!(t > -2147483648.0f && t < 2147483648.0f)) return INT_MIN;
    return (int)t;
}

// frand(-1, 1) (S/Smoke_solver3D.h:36-39: a + (b - a) U) with U from the
// counter hash: U = (randhash(seed * 0x9E3779B1 + key) >> 8) / 2^24.
__device__ inline float reseed_jitter(unsigned seed, unsigned key) {
    const unsigned r = randhash(seed * 0x9E3779B1u + key);
    const float u = (float)(r >> 8) * (1.0f / 16777216.0f);
    return -1.0f + 2.0f * u;
}

// Exclusive scan of a[0..n) in place, the total in a[n]; partial holds
// ceil(n / scan_chunk()) ints. Asynchronous on s.
size_t scan_chunk();
void scan_exclusive_i32(int* a, size_t n, int* partial, hipStream_t s);
// Grows P's selection scratch (pflag, ppartial) to n + 1 entries.
int ensure_selection_scratch(vflip_particles* P, size_t n);

}  // namespace vflip
