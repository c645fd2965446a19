// Kernels of the learned warm start: a 3D U-Net over channels-last BF16
// activations with FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16). Internal header of unet.hip.
//
// Layouts
//   activation  [z][y][x][C]  bf16, C padded to a multiple of 16 (pad = 0)
//   weights     [Cout][Kw]    bf16, k = tap*Cin + ci, tap = (dz+1)*9 + (dy+1)*3 + (dx+1);
//                             Kw = 27*Cin (28*Cin when Cin == 16: a zero tap pads K to 32s)
//   gradients   fp32, same shapes as what they differentiate
#pragma once

#include "common.h"

namespace vflip {
namespace unet {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Output tile of a convolution workgroup: 16 x 8 x 4 voxels, 32 rows of 16 x
// voxels. Wave w owns rows y = 4 (w & 1) + mb (mb = 0..3) of z-slice w >> 1, one
// row per 16-row M block of the matrix cores.
constexpr int kTX = 16, kTY = 8, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;  // halo tile 18 x 10 x 6
constexpr int kHVox = kHX * kHY * kHZ;
constexpr int kConvThreads = 512;  // 8 waves

// One input of a convolution: channels [c_lo, c_lo + C) of the (virtually
// concatenated) input come from x, read at (v >> up) on a grid of dims >> up.
struct ConvSrc {
    const bf16* x;
    int C;   // channels stored per voxel in x (padded)
    int up;  // 1: x lives on the half-resolution grid (nearest upsampling)
};

struct ConvArgs {
    ConvSrc src[2];
    int nsrc;
    int nx, ny, nz;  // output grid (and full-resolution input grid)
    int cin;         // total input channels (sum of src C)
    int cout;        // output channels (multiple of 16)
    int kw;          // weight row length
    const bf16* w;   // [cout][kw]
    const float* bias;
    int relu;
    bf16* out;       // [V][cout] bf16 (if not null)
    float* out32;    // [V][cout] fp32 (if not null), or [V] channel 0 only when out32_ch0
    int out32_ch0;
    int tiles_x, tiles_y, ntiles;
    int groups;  // output-channel groups: workgroup L computes channels of group L / ntiles
    int ksplit;  // > 1: the K loop is split over ksplit workgroups that write FP32 partials
    float* part; // [ksplit][V][cout] partials, summed by k_conv_splitk_reduce
    bf16* pool;  // with a.out: also the 2x2x2 means of the outputs, [V / 8][cout] (or nullptr)
};

// LDS layout of the halo tile of CK input channels: P = CK / 8 channel planes of
// 16 bytes per voxel, voxel (hx, hy, hz) at plane * PS + hz * Z + hy * R + hx.
// gfx950 serves a 16-byte LDS read 16 lanes at a time from 64 four-byte banks;
// it is free of bank conflicts when each 16-lane group reads 256 contiguous
// bytes and the four groups of the wave start at the same offset modulo 256
// bytes. A matrix-core A fragment is 16 consecutive x voxels (one group per
// channel plane, or per tap of a pair), so: PS is a multiple of 16 voxels, and
// for CK = 16, whose K steps pair two taps, Z = 191 makes the offsets of the
// paired taps (kTapPairs16) equal modulo 16 voxels.
template <int CK>
struct Halo {
    static constexpr int P = CK / 8;
    static constexpr int R = kHX;
    static constexpr int Z = CK == 16 ? 191 : kHX * kHY;
    static constexpr int PS = (kHZ * Z + 15) / 16 * 16;
    static constexpr int LDS = P * PS * 8;  // bf16 elements
    // 16-byte elements, in blocks of 64 (one wave's load): lane l of block b
    // takes part (l >> 4) % P of voxel b * 64 / P + (l & 15) + 16 ((l >> 4) / P)
    static constexpr int BLOCKS = (kHVox * P + 63) / 64;
    static constexpr int N = (BLOCKS * 64 + kConvThreads - 1) / kConvThreads;  // elements per thread
};

__device__ inline int tap_offset(int tap, int R, int Z) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    return dz * Z + dy * R + dx;
}

// K step st of a 16-channel chunk: taps (kTapPairs16 >> 10 st) & 31 and
// (>> 10 st + 5) & 31; tap 27 is the zero pad of the weights. Pairs:
// (0,10) (1,11) (2,3) (4,14) (5,6) (7,17) (8,27) (9,19) (12,22) (13,23)
// (15,25) (16,26) (18,27) (20,21) (24,27), four steps per 64-bit word.
__device__ inline void tap_pair16(int st, int& ta, int& tb) {
    constexpr unsigned long long W[4] = {
        (0ull | 10ull << 5) | (1ull | 11ull << 5) << 16 | (2ull | 3ull << 5) << 32 | (4ull | 14ull << 5) << 48,
        (5ull | 6ull << 5) | (7ull | 17ull << 5) << 16 | (8ull | 27ull << 5) << 32 | (9ull | 19ull << 5) << 48,
        (12ull | 22ull << 5) | (13ull | 23ull << 5) << 16 | (15ull | 25ull << 5) << 32 | (16ull | 26ull << 5) << 48,
        (18ull | 27ull << 5) | (20ull | 21ull << 5) << 16 | (24ull | 27ull << 5) << 32};
    const int q = st >> 2;
    const unsigned long long w = q == 0 ? W[0] : q == 1 ? W[1] : q == 2 ? W[2] : W[3];
    const unsigned v = (unsigned)(w >> ((st & 3) * 16));
    ta = v & 31;
    tb = (v >> 5) & 31;
}

// Global loads of G of this thread's halo elements (element threadIdx + (i0 + i) *
// 512), all issued before any is used; voxels outside the grid read as zero.
template <int CK, int G>
__device__ inline void load_halo(const ConvArgs& a, int c0, int x0, int y0, int z0, int i0, bf16x8 (&hr)[G]) {
    int s = 0, cl = c0;
    if (a.nsrc > 1 && c0 >= a.src[0].C) {
        s = 1;
        cl = c0 - a.src[0].C;
    }
    const ConvSrc src = a.src[s];
    const int sx = a.nx >> src.up, sy = a.ny >> src.up;
    constexpr int P = Halo<CK>::P;
    const int lane = threadIdx.x & 63;
    const int part = (lane >> 4) % P, vl = (lane & 15) + 16 * ((lane >> 4) / P);
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const int blk = (threadIdx.x + (i0 + i) * kConvThreads) >> 6;
        const int hv = blk * (64 / P) + vl;
        const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        bf16x8 val;
#pragma unroll
        for (int q = 0; q < 8; ++q) val[q] = (bf16)0.f;
        if (hv < kHVox && gx >= 0 && gy >= 0 && gz >= 0 && gx < a.nx && gy < a.ny && gz < a.nz) {
            const size_t v = ((size_t)(gz >> src.up) * sy + (gy >> src.up)) * sx + (gx >> src.up);
            val = *reinterpret_cast<const bf16x8*>(src.x + v * src.C + cl + part * 8);
        }
        hr[i] = val;
    }
}

template <int CK, int G>
__device__ inline void store_halo(const bf16x8 (&hr)[G], int i0, bf16* lds) {
    using H = Halo<CK>;
    const int lane = threadIdx.x & 63;
    const int part = (lane >> 4) % H::P, vl = (lane & 15) + 16 * ((lane >> 4) / H::P);
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const int blk = (threadIdx.x + (i0 + i) * kConvThreads) >> 6;
        const int hv = blk * (64 / H::P) + vl;
        if (hv < kHVox) {
            const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
            *reinterpret_cast<bf16x8*>(lds + (part * H::PS + hz * H::Z + hy * H::R + hx) * 8) = hr[i];
        }
    }
}

// Stages the whole halo tile into LDS, at most four loads per thread in flight.
template <int CK>
__device__ inline void stage_halo(const ConvArgs& a, int c0, int x0, int y0, int z0, bf16* lds) {
    constexpr int N = Halo<CK>::N, G = CK == 16 ? 1 : 3;
#pragma unroll 1
    for (int i0 = 0; i0 < N; i0 += G) {  // elements past the tile are skipped by load and store
        bf16x8 hr[G];
        load_halo<CK, G>(a, c0, x0, y0, z0, i0, hr);
        store_halo<CK, G>(hr, i0, lds);
    }
}

// The weights of K steps [lo, lo + n) of one channel chunk, staged into LDS
// beside the halo: 16-byte element e = ((s * NB + nb) * 4 + kg) * 16 + r16 (s =
// step - lo) holds the 8 k values of lane (r16, kg) of output block nb, so each
// 16-lane group of a B fragment reads 256 contiguous bytes.
template <int CK, int NB>
struct WChunk {
    static constexpr int STEPS = CK == 32 ? 27 : 15;
    static constexpr int E = STEPS * NB * 64;  // a whole chunk
    static constexpr int N = (E + kConvThreads - 1) / kConvThreads;  // elements per thread, at most
};

template <int CK, int NB>
__device__ inline void load_wchunk(const ConvArgs& a, int chunk, int lo, int n, int cbase,
                                   bf16x8 (&wr)[WChunk<CK, NB>::N]) {
#pragma unroll
    for (int i = 0; i < WChunk<CK, NB>::N; ++i) {
        const int e = threadIdx.x + i * kConvThreads;
        if (e < n * NB * 64) {
            const int st = lo + e / (NB * 64), rem = e % (NB * 64);
            const int nb = rem >> 6, kg = (rem >> 4) & 3, r16 = rem & 15;
            int tap = st, coff = kg * 8;
            if (CK == 16) {
                int ta, tb;
                tap_pair16(st, ta, tb);
                tap = (kg >> 1) ? tb : ta;
                coff = (kg & 1) * 8;
            }
            wr[i] = *reinterpret_cast<const bf16x8*>(a.w + (size_t)(cbase + nb * 16 + r16) * a.kw +
                                                     (size_t)tap * a.cin + chunk * CK + coff);
        }
    }
}

template <int CK, int NB>
__device__ inline void store_wchunk(const bf16x8 (&wr)[WChunk<CK, NB>::N], int n, bf16* ldw) {
#pragma unroll
    for (int i = 0; i < WChunk<CK, NB>::N; ++i) {
        const int e = threadIdx.x + i * kConvThreads;
        if (e < n * NB * 64) *reinterpret_cast<bf16x8*>(ldw + e * 8) = wr[i];
    }
}

// Epilogue pieces shared by the convolution kernels. D register r of lane
// (r16, kg) in M block mb, block nb holds voxel x = kg*4 + r of row (w & 1) * 4 +
// mb of z-slice w >> 1 of the tile, channel nb*16 + r16.

// The tile's BF16 outputs (bias, ReLU) into LDS as [voxel m][NB * 16], m =
// (zl * 8 + yl) * 16 + xl.
template <int NB>
__device__ inline void tile_to_lds(const ConvArgs& a, const f32x4 (&acc)[4][NB], const float (&bias)[NB], bf16* out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r16 = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = ((w >> 1) * 8 + (w & 1) * 4 + mb) * 16 + kg * 4 + r;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                float val = acc[mb][nb][r] + bias[nb];
                if (a.relu) val = val > 0.f ? val : 0.f;
                out[m * NB * 16 + nb * 16 + r16] = (bf16)val;
            }
        }
}

// The tile's outputs from LDS to a.out, 16-byte chunks of whole voxels.
template <int NB>
__device__ inline void lds_to_out(const ConvArgs& a, const bf16* out, int x0, int y0, int z0, int cbase) {
#pragma unroll
    for (int i = 0; i < 2 * NB; ++i) {
        const int j = threadIdx.x + i * kConvThreads;
        const int m = j / (2 * NB), part = j % (2 * NB);
        const int x = x0 + (m & 15), y = y0 + ((m >> 4) & 7), z = z0 + (m >> 7);
        if (x < a.nx && y < a.ny && z < a.nz) {
            const size_t v = ((size_t)z * a.ny + y) * a.nx + x;
            *reinterpret_cast<bf16x8*>(a.out + v * a.cout + cbase + part * 8) =
                *reinterpret_cast<const bf16x8*>(out + m * NB * 16 + part * 8);
        }
    }
}

// 2x2x2 means of the tile's BF16 outputs in LDS, each the FP32 sum of the 8 BF16
// values in k_avgpool's order times 0.125 (the same bits as k_avgpool on the
// stored outputs), two channels per thread. Written to a.pool (the half-
// resolution grid, a.cout channels), or (pl != nullptr) to LDS as [64 coarse
// voxels of the tile][NB * 16] for lds_pool_copy.
template <int NB>
__device__ inline void lds_pool(const ConvArgs& a, const bf16* out, int x0, int y0, int z0, int cbase, bf16* pl) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    constexpr int P = 8 * NB;  // channel pairs per voxel
    const int CX = a.nx / 2, CY = a.ny / 2, CZ = a.nz / 2;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int j = threadIdx.x + i * kConvThreads;  // 64 coarse voxels x P pairs
        const int cm = j / P, cp = j % P;
        const int ci = cm & 7, cj = (cm >> 3) & 3, ck = cm >> 5;  // coarse voxel in the tile's 8 x 4 x 2
        float s0 = 0.f, s1 = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int m = ((2 * ck + (q >> 2)) * 8 + (2 * cj + ((q >> 1) & 1))) * 16 + 2 * ci + (q & 1);
            const bf16x2 v = *reinterpret_cast<const bf16x2*>(out + m * NB * 16 + cp * 2);
            s0 += (float)v[0];
            s1 += (float)v[1];
        }
        bf16x2 o;
        o[0] = (bf16)(0.125f * s0);
        o[1] = (bf16)(0.125f * s1);
        if (pl) {
            *reinterpret_cast<bf16x2*>(pl + cm * NB * 16 + cp * 2) = o;
        } else {
            const int cx = x0 / 2 + ci, cy = y0 / 2 + cj, cz = z0 / 2 + ck;
            if (cx < CX && cy < CY && cz < CZ)
                *reinterpret_cast<bf16x2*>(a.pool + (((size_t)cz * CY + cy) * CX + cx) * a.cout + cbase + cp * 2) = o;
        }
    }
}

// The pooled values of a tile from LDS (lds_pool) to a.pool, 16-byte chunks.
template <int NB>
__device__ inline void lds_pool_copy(const ConvArgs& a, const bf16* pl, int x0, int y0, int z0, int cbase) {
    const int j = threadIdx.x;  // 64 coarse voxels x 2 NB 16-byte parts
    if (j >= 64 * 2 * NB) return;
    const int cm = j / (2 * NB), part = j % (2 * NB);
    const int cx = x0 / 2 + (cm & 7), cy = y0 / 2 + ((cm >> 3) & 3), cz = z0 / 2 + (cm >> 5);
    if (cx < a.nx / 2 && cy < a.ny / 2 && cz < a.nz / 2)
        *reinterpret_cast<bf16x8*>(a.pool + (((size_t)cz * (a.ny / 2) + cy) * (a.nx / 2) + cx) * a.cout + cbase +
                                   part * 8) = *reinterpret_cast<const bf16x8*>(pl + cm * NB * 16 + part * 8);
}

// FP32 channel 0 alone (a.out32_ch0; a.nx a multiple of 4): lanes r16 = 0 of
// block 0 hold it for 4 consecutive x. Written to out32, or (lds != nullptr)
// to LDS as [m] floats.
template <int NB>
__device__ inline void tile_ch0(const ConvArgs& a, const f32x4 (&acc)[4][NB], float bias0, int x0, int y0, int z0,
                                float* lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r16 = lane & 15, kg = lane >> 4;
    if (r16 != 0) return;
    const int z = z0 + (w >> 1);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        float4 o;
        o.x = acc[mb][0][0] + bias0;
        o.y = acc[mb][0][1] + bias0;
        o.z = acc[mb][0][2] + bias0;
        o.w = acc[mb][0][3] + bias0;
        if (a.relu) {
            o.x = fmaxf(o.x, 0.f);
            o.y = fmaxf(o.y, 0.f);
            o.z = fmaxf(o.z, 0.f);
            o.w = fmaxf(o.w, 0.f);
        }
        if (lds) {
            *reinterpret_cast<float4*>(lds + ((w >> 1) * 8 + (w & 1) * 4 + mb) * 16 + kg * 4) = o;
        } else {
            const int x = x0 + kg * 4, y = y0 + (w & 1) * 4 + mb;
            if (x < a.nx && y < a.ny && z < a.nz) *reinterpret_cast<float4*>(a.out32 + ((size_t)z * a.ny + y) * a.nx + x) = o;
        }
    }
}

// Channel 0 from LDS ([m] floats) to out32, 16 bytes (4 x voxels) per thread.
__device__ inline void lds_ch0_to_out32(const ConvArgs& a, const float* lds, int x0, int y0, int z0) {
    if (threadIdx.x >= 128) return;
    const int m = threadIdx.x * 4;
    const int x = x0 + (m & 15), y = y0 + ((m >> 4) & 7), z = z0 + (m >> 7);
    if (x < a.nx && y < a.ny && z < a.nz)
        *reinterpret_cast<float4*>(a.out32 + ((size_t)z * a.ny + y) * a.nx + x) =
            *reinterpret_cast<const float4*>(lds + m);
}

// LDS of a convolution workgroup (bf16 elements): the halo, plus the weights of
// the chunk's K steps where they fit (up to 3 output blocks: 16-channel chunks,
// and 32-channel chunks of split K loops, which run on small grids at one
// workgroup per CU). The compiler is held
// to the register budget of the waves per SIMD that this LDS allows (2 per
// workgroup, 160 KB per CU), unless the workgroup needs more registers.
__host__ __device__ constexpr bool conv_wlds(int ck, int nb, bool split) { return (ck == 16 || split) && nb <= 3; }
__host__ __device__ constexpr int conv_lds(int ck, int nb, bool split) {
    return (ck == 16 ? Halo<16>::LDS : Halo<32>::LDS) + (conv_wlds(ck, nb, split) ? (ck == 32 ? 27 : 15) * nb * 64 * 8 : 0);
}
// The outputs of a tile (512 voxels x 16 NB BF16) fit in the workgroup's LDS: an
// unsplit epilogue writes through LDS (16-byte stores, pooling).
__host__ __device__ constexpr bool conv_tile_in_lds(int ck, int nb, bool split) {
    return 512 * nb * 16 <= conv_lds(ck, nb, split);
}
template <int CK, int NB, bool SPLIT>
constexpr bool conv_wlds() { return conv_wlds(CK, NB, SPLIT); }
template <int CK, int NB, bool SPLIT>
constexpr int conv_lds() { return conv_lds(CK, NB, SPLIT); }
template <int CK, int NB, bool SPLIT>
constexpr int conv_waves_per_simd() {
    const int b = (160 * 1024) / (conv_lds<CK, NB, SPLIT>() * 2);
    const int w = 2 * (b < 1 ? 1 : (b > 4 ? 4 : b));
    // registers: accumulators, loads in flight, A fragments, addressing
    const int ring = conv_wlds<CK, NB, SPLIT>() ? 4 * (Halo<CK>::N + WChunk<CK, NB>::N) + 16
                                                : 4 * NB * (SPLIT ? (NB == 1 ? 2 : NB == 2 ? 4 : 1) : 3);
    const int cap = 512 / (16 * NB + ring + 56);
    return w < cap ? w : (cap < 2 ? 2 : cap);
}

// 3x3x3 convolution, implicit GEMM: M = 512 voxels of a 16 x 8 x 4 tile, N =
// cout, K = 27 * cin. Each K step of 32 is one tap x 32 channels (CK = 32, 27
// steps per chunk) or a pair of taps x 16 channels (CK = 16, 15 steps by
// tap_pair16, three of them with the zero pad tap). A workgroup computes NB blocks of 16 output
// channels, group (blockIdx / ntiles) % groups of the a.groups groups (cout =
// 16 NB groups); small grids split their channels over groups so that more
// workgroups share a layer. With ksplit > 1 workgroup slice blockIdx /
// (ntiles groups) takes a contiguous range of the K steps (channel chunk x
// step, the unsplit order) and writes its FP32 partial sums instead of the
// outputs; with ksplit 1 the loop is the unsplit one.
template <int CK, int NB, bool SPLIT>
__global__ __launch_bounds__(kConvThreads, (conv_waves_per_simd<CK, NB, SPLIT>())) void k_conv3(ConvArgs a) {
    using H = Halo<CK>;
    constexpr bool WLDS = conv_wlds<CK, NB, SPLIT>();  // the chunk's weights in LDS
    constexpr int LDS_ALL = conv_lds<CK, NB, SPLIT>();
    __shared__ __attribute__((aligned(16))) bf16 lds[LDS_ALL];
    const int L = blockIdx.x % a.ntiles, grp = (blockIdx.x / a.ntiles) % a.groups;
    const int ks = blockIdx.x / (a.ntiles * a.groups), nks = a.ksplit > 1 ? a.ksplit : 1;
    const int cbase = grp * NB * 16;
    int t = L;
    if (a.ntiles % 8 == 0) t = (L % 8) * (a.ntiles / 8) + L / 8;
    const int x0 = (t % a.tiles_x) * kTX, y0 = ((t / a.tiles_x) % a.tiles_y) * kTY;
    const int z0 = (t / (a.tiles_x * a.tiles_y)) * kTZ;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kg = lane >> 4;  // MFMA row/col within a block, k group of 8
    f32x4 acc[4][NB];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // LDS slot of tap (-1,-1,-1) of this lane's output voxel in M block 0, in its
    // channel plane; M block mb and tap t add mb * R + tap_offset(t) + Z + R + 1 >= 0,
    // compile-time offsets of the unrolled steps
    const int plane = CK == 32 ? kg : (kg & 1);
    const int hbase = plane * H::PS + (w >> 1) * H::Z + (w & 1) * 4 * H::R + r16;
    const int lane_hi = CK == 16 ? (kg >> 1) : 0;  // CK = 16: lanes of the second tap of the pair
    constexpr int STEPS = CK == 32 ? 27 : 15;
    const int total = ((a.cin + CK - 1) / CK) * STEPS;
    const int q0 = (int)((long long)ks * total / nks), q1 = (int)((long long)(ks + 1) * total / nks);
    // K step q: the taps of its two halves of k (the same tap for CK = 32; tap 27
    // is the zero pad of the weights, read at its partner's voxels)
    auto step_taps = [&](int q, int& ta, int& tb) {
        const int st = q % STEPS;
        ta = tb = st;
        if (CK == 16) tap_pair16(st, ta, tb);
    };
    // weights: a uniform base per tap plus one per-lane 32-bit byte offset (row
    // cbase + r16, k group, chunk; the second tap of a pair on lanes lane_hi)
    const unsigned wlane = ((unsigned)(cbase + r16) * (unsigned)a.kw + (CK == 32 ? kg * 8 : (kg & 1) * 8)) * 2u;
    auto load_b = [&](int q, bf16x8 (&bf)[NB]) {
        int ta, tb;
        step_taps(q, ta, tb);
        const char* base = reinterpret_cast<const char*>(a.w) + (size_t)ta * a.cin * 2;
        const unsigned off = wlane + (unsigned)((q / STEPS) * CK * 2) + (unsigned)(lane_hi * (tb - ta) * a.cin * 2);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
            bf[nb] = *reinterpret_cast<const bf16x8*>(base + (size_t)(off + (unsigned)(nb * 16 * a.kw * 2)));
    };
    // the A fragments of step st of a chunk for the four M blocks
    auto load_a = [&](int st, bf16x8 (&af)[4]) {
        int ta, tb;
        step_taps(st, ta, tb);
        const int oa = tap_offset(ta, H::R, H::Z) + H::Z + H::R + 1;
        const int ob = tap_offset(tb < 27 ? tb : ta, H::R, H::Z) + H::Z + H::R + 1;
        const bf16* p = lds + (hbase + lane_hi * (ob - oa)) * 8;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) af[mb] = *reinterpret_cast<const bf16x8*>(p + (oa + mb * H::R) * 8);
    };
    if constexpr (WLDS) {
        // Halo and weights of each chunk staged together (every load issued before
        // the first store), then the chunk's K steps from LDS alone: all of them,
        // unrolled, for a whole K loop; this slice's range for a split one.
        bf16* ldw = lds + H::LDS;
        auto step = [&](int st, int s) {
            bf16x8 af[4], bf[NB];
            load_a(st, af);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                bf[nb] = *reinterpret_cast<const bf16x8*>(ldw + (((s * NB + nb) * 4 + kg) * 16 + r16) * 8);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb], bf[nb], acc[mb][nb], 0, 0, 0);
        };
        const int c_hi = (q1 - 1) / STEPS;
        for (int c = q0 / STEPS; c <= c_hi; ++c) {
            const int lo = SPLIT ? max(q0 - c * STEPS, 0) : 0;
            const int hi = SPLIT ? min(q1 - c * STEPS, STEPS) : STEPS;
            __syncthreads();
            {
                bf16x8 hr[H::N], wr[WChunk<CK, NB>::N];
                load_halo<CK, H::N>(a, c * CK, x0, y0, z0, 0, hr);
                load_wchunk<CK, NB>(a, c, lo, hi - lo, cbase, wr);
                store_halo<CK, H::N>(hr, 0, lds);
                store_wchunk<CK, NB>(wr, hi - lo, ldw);
            }
            __syncthreads();
            if constexpr (!SPLIT) {
#pragma unroll
                for (int st = 0; st < STEPS; ++st) step(st, st);
            } else {
                for (int st = lo; st < hi; ++st) step(st, st - lo);
            }
        }
    } else if constexpr (!SPLIT) {
        // Whole chunks: the chunk's K steps unrolled, the weight fragments of step
        // st + 3 (of this chunk or the next) loaded while step st computes. 3
        // divides the steps per chunk, so ring slot st % 3 holds step st in every
        // chunk; past the last chunk the prefetch repeats it (unused).
        constexpr int D = 3;
        static_assert(STEPS % D == 0, "ring slots");
        const int nch = (a.cin + CK - 1) / CK;
        bf16x8 ring[D][NB];
#pragma unroll
        for (int d = 0; d < D; ++d) load_b(d, ring[d]);
        for (int c = 0; c < nch; ++c) {
            __syncthreads();
            stage_halo<CK>(a, c * CK, x0, y0, z0, lds);
            __syncthreads();
#pragma unroll
            for (int st = 0; st < STEPS; ++st) {
                const int q = c * STEPS + st;
                bf16x8 af[4];
                load_a(st, af);
#pragma unroll
                for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[mb][nb] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb], ring[st % D][nb], acc[mb][nb], 0, 0, 0);
                load_b(min(q + D, nch * STEPS - 1), ring[st % D]);
            }
        }
    } else {
        constexpr int D = NB == 1 ? 2 : (NB == 2 ? 4 : 1);
        // Every step issues exactly one prefetch (past the end: the last step
        // again, unused).
        bf16x8 ring[D][NB];
#pragma unroll
        for (int d = 0; d < D; ++d) load_b(min(q0 + d, q1 - 1), ring[d]);
        int staged = -1;
        for (int qb = q0; qb < q1; qb += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int q = qb + d;
                if (q >= q1) break;  // uniform over the workgroup
                const int c0 = (q / STEPS) * CK;
                if (c0 != staged) {  // uniform over the workgroup
                    __syncthreads();
                    stage_halo<CK>(a, c0, x0, y0, z0, lds);
                    __syncthreads();
                    staged = c0;
                }
                bf16x8 af[4];
                load_a(q % STEPS, af);
#pragma unroll
                for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb], ring[d][nb], acc[mb][nb], 0, 0, 0);
                load_b(min(q + D, q1 - 1), ring[d]);
            }
        }
    }
    // epilogue: D reg r of lane -> voxel x = kg*4 + r of the M block's row, channel nb*16 + r16.
    // The bias is read before any store: the FP32 output could alias it as far as
    // the compiler knows, and a load issued after stores waits for all of them.
    float bias[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bias[nb] = a.bias && nks == 1 ? a.bias[cbase + nb * 16 + r16] : 0.f;
    if constexpr (!SPLIT && conv_tile_in_lds(CK, NB, false)) {
        if (a.out && !a.out32) {  // uniform: BF16 output only
            // the tile's outputs gathered in LDS, written as 16-byte chunks of whole voxels
            __syncthreads();  // every wave is done with the halo and weights
            tile_to_lds<NB>(a, acc, bias, lds);
            __syncthreads();
            lds_to_out<NB>(a, lds, x0, y0, z0, cbase);
            if (a.pool) lds_pool<NB>(a, lds, x0, y0, z0, cbase, nullptr);
            return;
        }
    }
    const int z = z0 + (w >> 1);
    if (!SPLIT && a.out32 && a.out32_ch0 && !a.out && (a.nx & 3) == 0) {  // uniform
        if (cbase == 0) tile_ch0<NB>(a, acc, bias[0], x0, y0, z0, nullptr);
        return;
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int x = x0 + kg * 4 + r, y = y0 + (w & 1) * 4 + mb;
            if (x >= a.nx || y >= a.ny || z >= a.nz) continue;
            const size_t v = ((size_t)z * a.ny + y) * a.nx + x;
            if (nks > 1) {
                const size_t V = (size_t)a.nx * a.ny * a.nz;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    a.part[((size_t)ks * V + v) * a.cout + cbase + nb * 16 + r16] = acc[mb][nb][r];
                continue;
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = cbase + nb * 16 + r16;
                float val = acc[mb][nb][r];
                if (a.bias) val += bias[nb];
                if (a.relu) val = val > 0.f ? val : 0.f;
                if (a.out) a.out[v * a.cout + co] = (bf16)val;
                if (a.out32) {
                    if (!a.out32_ch0) a.out32[v * a.cout + co] = val;
                    else if (co == 0) a.out32[v] = val;
                }
            }
        }
}

// k_conv_splitk_reduce for a BF16 output that also feeds a 2x2x2 average pool
// (a.pool; even grid dims): one thread per (coarse voxel, channel) reduces its 8
// fine voxels in k_avgpool's order and writes them and their mean, the same
// bits as k_conv_splitk_reduce followed by k_avgpool.
__global__ __launch_bounds__(256) void k_conv_splitk_reduce_pool(ConvArgs a) {
    const int CX = a.nx / 2, CY = a.ny / 2, CZ = a.nz / 2;
    const size_t V = (size_t)a.nx * a.ny * a.nz, nf = V * a.cout, n = (size_t)CX * CY * CZ * a.cout;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const int co = (int)(e % (unsigned)a.cout);
        const size_t cv = e / (unsigned)a.cout;
        const int i = (int)(cv % CX), j = (int)((cv / CX) % CY), k = (int)(cv / ((size_t)CX * CY));
        const float b = a.bias ? a.bias[co] : 0.f;
        size_t fe[8];
        bf16 o[8];
        float sum = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {  // every load before the first store (a load waits for earlier stores)
            const size_t f = ((size_t)(2 * k + (q >> 2)) * a.ny + (2 * j + ((q >> 1) & 1))) * a.nx + (2 * i + (q & 1));
            fe[q] = f * a.cout + co;
            float val = a.part[fe[q]];
            for (int sl = 1; sl < a.ksplit; ++sl) val += a.part[(size_t)sl * nf + fe[q]];
            val += b;
            if (a.relu) val = val > 0.f ? val : 0.f;
            o[q] = (bf16)val;
            sum += (float)o[q];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) a.out[fe[q]] = o[q];
        a.pool[e] = (bf16)(0.125f * sum);
    }
}

// Persistent form of k_conv3 for whole K loops of 16-channel chunks into one
// block of 16 outputs on large grids (a.groups 1, BF16 output or FP32 channel
// 0). gridDim.x workgroups, as many as fit on the GPU at once, walk tiles
// blockIdx.x, blockIdx.x + gridDim.x, ..., each over its chunks. The halo (and,
// with several chunks, the weights) of the next (tile, chunk) is loaded into
// registers while the current one computes and is stored into LDS after it;
// with one chunk the weights are staged once. A tile's outputs wait in LDS and
// are written to global memory after the next halo has arrived, so that the
// wait for that halo never includes their stores. Per tile the arithmetic and
// K order are k_conv3's: the outputs are the same bits. The tile's 2x2x2 means
// (a.pool) are formed from its outputs in LDS and written with them.
constexpr int kOutLds = 512 * 16;  // bf16: one tile of 16 BF16 outputs, or 512 FP32 channel-0 values
__global__ __launch_bounds__(kConvThreads, 4) void k_conv3_tiles(ConvArgs a) {
    constexpr int CK = 16, NB = 1, STEPS = 15;
    using H = Halo<CK>;
    using W = WChunk<CK, NB>;
    __shared__ __attribute__((aligned(16))) bf16 lds[H::LDS + W::E * 8 + kOutLds + 64 * 16];
    bf16* ldw = lds + H::LDS;
    bf16* ldo = ldw + W::E * 8;
    bf16* ldp = ldo + kOutLds;  // the pending tile's pooled outputs
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kg = lane >> 4, lane_hi = kg >> 1;
    const int hbase = (kg & 1) * H::PS + (w >> 1) * H::Z + (w & 1) * 4 * H::R + r16;
    const int nch = (a.cin + CK - 1) / CK;
    if ((int)blockIdx.x >= a.ntiles) return;  // uniform
    const int nseq = ((a.ntiles - 1 - (int)blockIdx.x) / (int)gridDim.x + 1) * nch;
    const bool ch0 = a.out32 != nullptr;  // FP32 channel 0 instead of BF16 outputs
    // sequence item k: tile origin and chunk
    auto at = [&](int k, int& x0, int& y0, int& z0, int& c) {
        c = k % nch;
        const int L = (int)blockIdx.x + (k / nch) * (int)gridDim.x;
        int t = L;
        if (a.ntiles % 8 == 0) t = (L % 8) * (a.ntiles / 8) + L / 8;
        x0 = (t % a.tiles_x) * kTX;
        y0 = ((t / a.tiles_x) % a.tiles_y) * kTY;
        z0 = (t / (a.tiles_x * a.tiles_y)) * kTZ;
    };
    // bias before any store (a load after stores would wait for them)
    const float bias = a.bias ? a.bias[r16] : 0.f;
    const float bias0 = __shfl(bias, 0);
    bf16x8 hr[H::N], wr[W::N];
    {
        int x0, y0, z0, c;
        at(0, x0, y0, z0, c);
        load_halo<CK, H::N>(a, c * CK, x0, y0, z0, 0, hr);
        load_wchunk<CK, NB>(a, c, 0, STEPS, 0, wr);
    }
    f32x4 acc[4][NB];
    int px0 = 0, py0 = 0, pz0 = 0;
    bool pending = false;
    for (int k = 0; k < nseq; ++k) {
        int x0, y0, z0, c;
        at(k, x0, y0, z0, c);
        __syncthreads();  // the previous item's K loop and the pending outputs' LDS reads are done
        store_halo<CK, H::N>(hr, 0, lds);
        if (nch > 1 || k == 0) store_wchunk<CK, NB>(wr, STEPS, ldw);
        __syncthreads();
        if (k + 1 < nseq) {
            int x1, y1, z1, c1;
            at(k + 1, x1, y1, z1, c1);
            load_halo<CK, H::N>(a, c1 * CK, x1, y1, z1, 0, hr);
            if (nch > 1) load_wchunk<CK, NB>(a, c1, 0, STEPS, 0, wr);
        }
        if (pending) {
            if (ch0) {
                lds_ch0_to_out32(a, reinterpret_cast<const float*>(ldo), px0, py0, pz0);
            } else {
                lds_to_out<NB>(a, ldo, px0, py0, pz0, 0);
                if (a.pool) lds_pool_copy<NB>(a, ldp, px0, py0, pz0, 0);
            }
            pending = false;
        }
        if (c == 0) {
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) acc[mb][0] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int st = 0; st < STEPS; ++st) {
            int ta, tb;
            tap_pair16(st, ta, tb);
            const int oa = tap_offset(ta, H::R, H::Z) + H::Z + H::R + 1;
            const int ob = tap_offset(tb < 27 ? tb : ta, H::R, H::Z) + H::Z + H::R + 1;
            const bf16* p = lds + (hbase + lane_hi * (ob - oa)) * 8;
            const bf16x8 bf = *reinterpret_cast<const bf16x8*>(ldw + ((st * 4 + kg) * 16 + r16) * 8);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const bf16x8 af = *reinterpret_cast<const bf16x8*>(p + (oa + mb * H::R) * 8);
                acc[mb][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf, acc[mb][0], 0, 0, 0);
            }
            if (st % 2 == 1) __builtin_amdgcn_sched_barrier(0);  // bounds the loads issued ahead: no spills
        }
        if (c == nch - 1) {
            __syncthreads();  // every wave has read the pending outputs
            if (ch0) {
                tile_ch0<NB>(a, acc, bias0, x0, y0, z0, reinterpret_cast<float*>(ldo));
            } else {
                const float b1[1] = {bias};
                tile_to_lds<NB>(a, acc, b1, ldo);
                if (a.pool) {
                    __syncthreads();
                    lds_pool<NB>(a, ldo, x0, y0, z0, 0, ldp);
                }
            }
            px0 = x0;
            py0 = y0;
            pz0 = z0;
            pending = true;
        }
    }
    __syncthreads();
    if (pending) {
        if (ch0) {
            lds_ch0_to_out32(a, reinterpret_cast<const float*>(ldo), px0, py0, pz0);
        } else {
            lds_to_out<NB>(a, ldo, px0, py0, pz0, 0);
            if (a.pool) lds_pool_copy<NB>(a, ldp, px0, py0, pz0, 0);
        }
    }
}

// Sum of a split convolution's partials in slice order, then the epilogue of k_conv3.
__global__ __launch_bounds__(256) void k_conv_splitk_reduce(ConvArgs a) {
    const size_t V = (size_t)a.nx * a.ny * a.nz, n = V * a.cout;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const int co = (int)(e % (unsigned)a.cout);
        const size_t v = e / (unsigned)a.cout;
        float val = a.part[e];
        for (int s = 1; s < a.ksplit; ++s) val += a.part[(size_t)s * n + e];
        if (a.bias) val += a.bias[co];
        if (a.relu) val = val > 0.f ? val : 0.f;
        if (a.out) a.out[e] = (bf16)val;
        if (a.out32) {
            if (!a.out32_ch0) a.out32[e] = val;
            else if (co == 0) a.out32[v] = val;
        }
    }
}

// Weight gradient dW[co][tap*cin + ci] += sum_v g[v][co] * x[v + tap][ci]:
// MFMA with M = 16 output channels, N = 16 input channels of one tap, K = 32
// voxels per step. A workgroup (4 waves) owns one (tap, ci block) and a chunk of
// voxels; each wave reduces a quarter of it, the waves' partials are summed in
// LDS and added to dW with one atomic per entry.
struct WgradArgs {
    ConvSrc src[2];
    int nsrc;
    int nx, ny, nz;
    int cin, cout, kw;
    const bf16* g;  // [V][cout] gradient at the convolution output (after the ReLU mask)
    float* dw;      // [cout][kw]
    int chunk;      // voxels per workgroup (multiple of 128)
};

template <int NCB>
__global__ __launch_bounds__(256) void k_conv3_wgrad(WgradArgs a) {
    __shared__ float red[4][NCB][16][16];
    const int ncib = a.cin / 16;
    const int tap = blockIdx.x / ncib, cib = blockIdx.x % ncib;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    int ci0 = cib * 16, s = 0;
    if (a.nsrc > 1 && ci0 >= a.src[0].C) {
        s = 1;
        ci0 -= a.src[0].C;
    }
    const ConvSrc src = a.src[s];
    const int sx = a.nx >> src.up, sy = a.ny >> src.up;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kg = lane >> 4;
    const size_t V = (size_t)a.nx * a.ny * a.nz;
    const size_t v_begin = (size_t)blockIdx.y * a.chunk;
    const size_t v_end = min(V, v_begin + a.chunk);
    f32x4 acc[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (size_t vb = v_begin + (size_t)w * 32; vb < v_end; vb += 128) {
        bf16x8 bfrag;  // B[k = voxel][n = ci]: x at (voxel + tap), channel ci0 + r16
        bf16x8 afrag[NCB];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const size_t v = vb + kg * 8 + q;
            bf16 xv = (bf16)0.f;
            if (v < v_end) {
                const int x = (int)(v % a.nx), y = (int)((v / a.nx) % a.ny), z = (int)(v / ((size_t)a.nx * a.ny));
                const int gx = x + dx, gy = y + dy, gz = z + dz;
                if (gx >= 0 && gy >= 0 && gz >= 0 && gx < a.nx && gy < a.ny && gz < a.nz) {
                    const size_t sv = ((size_t)(gz >> src.up) * sy + (gy >> src.up)) * sx + (gx >> src.up);
                    xv = src.x[sv * src.C + ci0 + r16];
                }
            }
            bfrag[q] = xv;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) afrag[cb][q] = v < v_end ? a.g[v * a.cout + cb * 16 + r16] : (bf16)0.f;
        }
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag[cb], bfrag, acc[cb], 0, 0, 0);
    }
    // D: row (co within block) kg*4 + r, column (ci within block) r16
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[w][cb][kg * 4 + r][r16] = acc[cb][r];
    __syncthreads();
    for (int e = threadIdx.x; e < NCB * 256; e += 256) {
        const int cb = e / 256, row = (e / 16) % 16, col = e % 16;
        const float sum = red[0][cb][row][col] + red[1][cb][row][col] + red[2][cb][row][col] + red[3][cb][row][col];
        if (sum != 0.f)
            unsafeAtomicAdd(&a.dw[(size_t)(cb * 16 + row) * a.kw + (size_t)tap * a.cin + cib * 16 + col], sum);
    }
}

// db[co] += sum over voxels of g[v][co] (g bf16 after the ReLU mask).
__global__ __launch_bounds__(256) void k_bias_grad(const bf16* __restrict__ g, size_t V, int cout, float* __restrict__ db) {
    __shared__ float part[256];
    const int co = threadIdx.x % cout, lanes = 256 / cout * cout;
    float s = 0.f;
    if (threadIdx.x < lanes)
        for (size_t v = blockIdx.x * (size_t)(256 / cout) + threadIdx.x / cout; v < V; v += (size_t)gridDim.x * (256 / cout))
            s += (float)g[v * cout + co];
    part[threadIdx.x] = threadIdx.x < lanes ? s : 0.f;
    __syncthreads();
    if (threadIdx.x < cout) {
        float t = 0.f;
        for (int q = threadIdx.x; q < lanes; q += cout) t += part[q];
        unsafeAtomicAdd(&db[threadIdx.x], t);
    }
}

// 2x2x2 average pooling, channels-last bf16 (fine dims even).
__global__ __launch_bounds__(256) void k_avgpool(const bf16* __restrict__ x, int nx, int ny, int nz, int C,
                                                 bf16* __restrict__ y) {
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    const size_t n = (size_t)cx * cy * cz * C;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const int c = (int)(e % C);
        const size_t v = e / C;
        const int i = (int)(v % cx), j = (int)((v / cx) % cy), k = (int)(v / ((size_t)cx * cy));
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const size_t f = ((size_t)(2 * k + (q >> 2)) * ny + (2 * j + ((q >> 1) & 1))) * nx + (2 * i + (q & 1));
            s += (float)x[f * C + c];
        }
        y[e] = (bf16)(0.125f * s);
    }
}

// gx_fine[v][c] += 0.125 * gy_coarse[v/2][c] (fp32).
__global__ __launch_bounds__(256) void k_avgpool_bwd(const float* __restrict__ gy, int nx, int ny, int nz, int C,
                                                     float* __restrict__ gx) {
    const size_t n = (size_t)nx * ny * nz * C;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const int c = (int)(e % C);
        const size_t v = e / C;
        const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((size_t)nx * ny));
        const size_t cv = ((size_t)(k / 2) * (ny / 2) + j / 2) * (nx / 2) + i / 2;
        gx[e] += 0.125f * gy[cv * C + c];
    }
}

// Splits the gradient of a virtual concat [up(coarse) C0 | skip C1] (fp32
// [V][C0+C1]): the upsampled part is summed over each 2x2x2 block into the
// coarse gradient (set), the skip part is added to the skip gradient.
__global__ __launch_bounds__(256) void k_concat_bwd_up(const float* __restrict__ g, int nx, int ny, int nz, int C0,
                                                       int C1, float* __restrict__ gcoarse) {
    const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
    const size_t n = (size_t)cx * cy * cz * C0;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const int c = (int)(e % C0);
        const size_t v = e / C0;
        const int i = (int)(v % cx), j = (int)((v / cx) % cy), k = (int)(v / ((size_t)cx * cy));
        float s = 0.f;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const size_t f = ((size_t)(2 * k + (q >> 2)) * ny + (2 * j + ((q >> 1) & 1))) * nx + (2 * i + (q & 1));
            s += g[f * (C0 + C1) + c];
        }
        gcoarse[e] = s;
    }
}

__global__ __launch_bounds__(256) void k_concat_bwd_skip(const float* __restrict__ g, size_t V, int C0, int C1,
                                                         float* __restrict__ gskip) {
    const size_t n = V * C1;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const size_t v = e / C1;
        gskip[e] += g[v * (C0 + C1) + C0 + (e % C1)];
    }
}

// g_pre = g * (y > 0) as bf16 (y: the layer's ReLU output); relu = 0 passes g.
__global__ __launch_bounds__(256) void k_relu_bwd(const float* __restrict__ g, const bf16* __restrict__ y, size_t n,
                                                  int relu, bf16* __restrict__ gp) {
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const float v = g[e];
        gp[e] = (bf16)((!relu || (float)y[e] > 0.f) ? v : 0.f);
    }
}

// Scale of the automatic normalisation: s = max|rhs| on the device (no host
// round trip); the network sees rhs / s and predicts pressure / s.
__device__ inline float inv_of(const float* s_dev, float explicit_inv) {
    if (!s_dev) return explicit_inv;
    const float s = s_dev[0];
    return s > 0.f ? 1.0f / s : 0.f;
}

// *out = max |x| (out zeroed beforehand); non-negative floats order as their bits.
__global__ __launch_bounds__(256) void k_absmax_to(const float* __restrict__ x, size_t n, float* __restrict__ out) {
    float m = 0.f;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) m = fmaxf(m, fabsf(x[e]));
    m = block_max(m);
    if (threadIdx.x == 0) atomicMax(reinterpret_cast<int*>(out), __float_as_int(m));
}

__global__ void k_zero1(float* a) { a[0] = 0.f; }

// Network input: channels 0..2 = rhs * inv_scale, solid, liquid; 3..15 = 0.
// inv_scale = 1/s_dev[0] when s_dev is given.
__global__ __launch_bounds__(256) void k_pack_input(const float* __restrict__ rhs, const float* __restrict__ solid,
                                                    const float* __restrict__ liquid, size_t V, float inv_scale_arg,
                                                    const float* __restrict__ s_dev, bf16* __restrict__ x) {
    const float inv_scale = inv_of(s_dev, inv_scale_arg);
    for (size_t v = blockIdx.x * (size_t)256 + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        bf16x8 lo, hi;
#pragma unroll
        for (int q = 0; q < 8; ++q) lo[q] = hi[q] = (bf16)0.f;
        lo[0] = (bf16)(rhs[v] * inv_scale);
        lo[1] = (bf16)(solid ? solid[v] : 0.f);
        lo[2] = (bf16)(liquid ? liquid[v] : 1.f);
        *reinterpret_cast<bf16x8*>(x + v * 16) = lo;
        *reinterpret_cast<bf16x8*>(x + v * 16 + 8) = hi;
    }
}

// MSE loss 0.5/V * sum (y - t)^2 (y: the prediction, [V]); its gradient is
// written to g[V] and the loss partials are added to *loss.
// The target is scaled by 1/(s_dev[0] * gain) when s_dev is given (normalised pressure).
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t, size_t V,
                                             const float* __restrict__ s_dev, float gain, float* __restrict__ g,
                                             double* __restrict__ loss) {
    double part = 0.0;
    const float inv = 1.0f / (float)V;
    const float tm = inv_of(s_dev, 1.0f) / (s_dev ? gain : 1.0f);
    for (size_t v = blockIdx.x * (size_t)256 + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        const float d = y[v] - t[v] * tm;
        part += 0.5 * (double)d * d / (double)V;
        g[v] = d * inv;
    }
    part = block_sum(part);
    if (threadIdx.x == 0) atomicAdd(loss, part);
}

// Energy-norm loss of a warm start. With x = target / (s * gain) (normalised
// pressure) and e = pred - x: loss = e'Ae / x'Ax, gradient 2 Ae / x'Ax, A the
// sample's 7-point operator (diag and the couplings to +x, +y, +z).
struct Stencil {
    const float *d, *ax, *ay, *az;
    int nx, ny, nz;
};

__device__ inline float stencil_apply(const Stencil& A, const float* x, size_t c, int i, int j, int k) {
    const size_t sx = 1, sy = (size_t)A.nx, sz = (size_t)A.nx * A.ny;
    float y = A.d[c] * x[c];
    if (i + 1 < A.nx) y += A.ax[c] * x[c + sx];
    if (i > 0) y += A.ax[c - sx] * x[c - sx];
    if (j + 1 < A.ny) y += A.ay[c] * x[c + sy];
    if (j > 0) y += A.ay[c - sy] * x[c - sy];
    if (k + 1 < A.nz) y += A.az[c] * x[c + sz];
    if (k > 0) y += A.az[c - sz] * x[c - sz];
    return y;
}

// e = pred - target * tm, xn = target * tm (tm = 1/(s * gain))
__global__ __launch_bounds__(256) void k_energy_prep(const float* __restrict__ pred, const float* __restrict__ t, size_t V,
                                                     const float* __restrict__ s_dev, float gain, float* __restrict__ e,
                                                     float* __restrict__ xn, double* __restrict__ acc) {
    const float tm = inv_of(s_dev, 1.0f) / gain;
    if (blockIdx.x == 0 && threadIdx.x < 2) acc[threadIdx.x] = 0.0;
    for (size_t v = blockIdx.x * (size_t)256 + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        const float x = t[v] * tm;
        xn[v] = x;
        e[v] = pred[v] - x;
    }
}

// acc[slot] += x'Ax; Ax written to y when y is given.
__global__ __launch_bounds__(256) void k_stencil_quad(Stencil A, const float* __restrict__ x, float* __restrict__ y,
                                                      double* __restrict__ acc, int slot) {
    const size_t V = (size_t)A.nx * A.ny * A.nz;
    double part = 0.0;
    for (size_t c = blockIdx.x * (size_t)256 + threadIdx.x; c < V; c += (size_t)gridDim.x * 256) {
        const int i = (int)(c % A.nx), j = (int)((c / A.nx) % A.ny), k = (int)(c / ((size_t)A.nx * A.ny));
        const float ax = stencil_apply(A, x, c, i, j, k);
        if (y) y[c] = ax;
        part += (double)x[c] * ax;
    }
    part = block_sum(part);
    if (threadIdx.x == 0) atomicAdd(&acc[slot], part);
}

// g = 2 Ae / x'Ax; loss = e'Ae / x'Ax.
__global__ __launch_bounds__(256) void k_energy_grad(const float* __restrict__ ae, size_t V,
                                                     const double* __restrict__ acc, float* __restrict__ g,
                                                     double* __restrict__ loss) {
    const double den = acc[1] > 0.0 ? acc[1] : 1.0;
    const float sc = (float)(2.0 / den);
    if (blockIdx.x == 0 && threadIdx.x == 0) *loss = acc[0] / den;
    for (size_t v = blockIdx.x * (size_t)256 + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) g[v] = ae[v] * sc;
}

// Gradient of a single-channel output [V] spread to the padded [V][16] layout
// as bf16 (channels 1..15 zero): the output layer has no ReLU.
__global__ __launch_bounds__(256) void k_grad_ch0(const float* __restrict__ g, size_t V, bf16* __restrict__ gp) {
    for (size_t v = blockIdx.x * (size_t)256 + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        bf16x8 lo, hi;
#pragma unroll
        for (int q = 0; q < 8; ++q) lo[q] = hi[q] = (bf16)0.f;
        lo[0] = (bf16)g[v];
        *reinterpret_cast<bf16x8*>(gp + v * 16) = lo;
        *reinterpret_cast<bf16x8*>(gp + v * 16 + 8) = hi;
    }
}

// Adam on the fp32 master copy; refreshes the bf16 forward weights.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2, float gscale, bf16* __restrict__ pb) {
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const float gr = g[e] * gscale;
        const float mm = b1 * m[e] + (1.f - b1) * gr;
        const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
        m[e] = mm;
        v[e] = vv;
        const float upd = lr * (mm * c1) / (sqrtf(vv * c2) + eps);
        const float np = p[e] - upd;
        p[e] = np;
        if (pb) pb[e] = (bf16)np;
    }
}

// Backward-data weights: wb[ci][tap' * cout + co] = w[co][tap * cin + ci] with
// tap' = 26 - tap (the transposed convolution is a convolution with the
// spatially flipped, channel-transposed kernel).
__global__ __launch_bounds__(256) void k_flip_weights(const float* __restrict__ w, int cin, int cout, int kw, int kwb,
                                                      bf16* __restrict__ wb) {
    const size_t n = (size_t)cin * kwb;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) {
        const int ci = (int)(e / kwb), kb = (int)(e % kwb);
        const int tb = kb / cout, co = kb % cout;
        float val = 0.f;
        if (tb < 27) val = w[(size_t)co * kw + (size_t)(26 - tb) * cin + ci];
        wb[e] = (bf16)val;
    }
}

// out = y * s_dev[0] * gain (or y when s_dev is null).
__global__ __launch_bounds__(256) void k_scale_copy(const float* __restrict__ y, size_t n, const float* __restrict__ s_dev,
                                                    float gain, float* __restrict__ out) {
    const float m = s_dev ? s_dev[0] * gain : 1.0f;
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) out[e] = y[e] * m;
}

// Solve grid -> network grid by an integer factor f: rhs averaged over each
// f^3 block; solid = fraction of cells outside the system (diagonal 0, the
// solid and Dirichlet cells); liquid = 1 - solid.
// 3D launch: x over coarse i (256 a block), y over coarse j, z over coarse k.
// vec4: f == 4, nx % 4 == 0 and rhs/diag 16-byte aligned (checked by the caller).
__global__ __launch_bounds__(256) void k_restrict_inputs(const float* __restrict__ rhs, const float* __restrict__ diag,
                                                         int nx, int ny, int nz, int f, int vec4,
                                                         float* __restrict__ crhs, float* __restrict__ csolid,
                                                         float* __restrict__ cliquid) {
    const int cx = nx / f, cy = ny / f;
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= cx) return;
    const float inv = 1.0f / (float)(f * f * f);
    float sr = 0.f, ss = 0.f;
    for (int dk = 0; dk < f; ++dk)
        for (int dj = 0; dj < f; ++dj) {
            const size_t row = ((size_t)(k * f + dk) * ny + (j * f + dj)) * nx + (size_t)i * f;
            if (vec4) {
                // the four fine values of a lane are one aligned float4 (same sum order)
                const float4 r4 = *reinterpret_cast<const float4*>(rhs + row);
                sr += r4.x; sr += r4.y; sr += r4.z; sr += r4.w;
                if (diag) {
                    const float4 d4 = *reinterpret_cast<const float4*>(diag + row);
                    ss += d4.x == 0.f ? 1.f : 0.f; ss += d4.y == 0.f ? 1.f : 0.f;
                    ss += d4.z == 0.f ? 1.f : 0.f; ss += d4.w == 0.f ? 1.f : 0.f;
                }
                continue;
            }
            for (int di = 0; di < f; ++di) {
                sr += rhs[row + di];
                ss += (diag && diag[row + di] == 0.f) ? 1.f : 0.f;
            }
        }
    const size_t c = ((size_t)k * cy + j) * cx + i;
    crhs[c] = sr * inv;
    csolid[c] = ss * inv;
    cliquid[c] = 1.f - ss * inv;
}

// Network grid -> solve grid: trilinear interpolation of cell-centred values
// (fine cell i sits at coarse coordinate (i + 0.5)/f - 0.5, clamped), times
// s_dev[0] * gain; cells outside the system (diag 0) get 0.
// 3D launch: x over fine i (256 a block), y over fine j, z over fine k.
__global__ __launch_bounds__(256) void k_prolong_cells(const float* __restrict__ cp, int cx, int cy, int cz, int f,
                                                       const float* __restrict__ s_dev, float gain,
                                                       const float* __restrict__ diag, float* __restrict__ x) {
    const int nx = cx * f, ny = cy * f;
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= nx) return;
    const size_t v = ((size_t)k * ny + j) * nx + i;
    const float m = s_dev ? s_dev[0] * gain : 1.0f;
    const float invf = 1.0f / (float)f;
    {
        if (diag && diag[v] == 0.f) {
            x[v] = 0.f;
            return;
        }
        const float gx = fminf(fmaxf((i + 0.5f) * invf - 0.5f, 0.f), (float)(cx - 1));
        const float gy = fminf(fmaxf((j + 0.5f) * invf - 0.5f, 0.f), (float)(cy - 1));
        const float gz = fminf(fmaxf((k + 0.5f) * invf - 0.5f, 0.f), (float)(cz - 1));
        const int i0 = (int)gx, j0 = (int)gy, k0 = (int)gz;
        const int i1 = min(i0 + 1, cx - 1), j1 = min(j0 + 1, cy - 1), k1 = min(k0 + 1, cz - 1);
        const float fx = gx - i0, fy = gy - j0, fz = gz - k0;
        auto at = [&](int a, int b, int e) { return cp[((size_t)e * cy + b) * cx + a]; };
        const float c00 = at(i0, j0, k0) + fx * (at(i1, j0, k0) - at(i0, j0, k0));
        const float c10 = at(i0, j1, k0) + fx * (at(i1, j1, k0) - at(i0, j1, k0));
        const float c01 = at(i0, j0, k1) + fx * (at(i1, j0, k1) - at(i0, j0, k1));
        const float c11 = at(i0, j1, k1) + fx * (at(i1, j1, k1) - at(i0, j1, k1));
        const float c0 = c00 + fy * (c10 - c00), c1 = c01 + fy * (c11 - c01);
        x[v] = (c0 + fz * (c1 - c0)) * m;
    }
}

// k_prolong_cells added into an FP64 guess: g += the prolonged value (same
// interpolation and scale; cells with diag 0 are left as they are).
__global__ __launch_bounds__(256) void k_prolong_add64(const float* __restrict__ cp, int cx, int cy, int cz, int f,
                                                       const float* __restrict__ s_dev, float gain,
                                                       const float* __restrict__ diag, double* __restrict__ g) {
    const int nx = cx * f, ny = cy * f;
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= nx) return;
    const size_t v = ((size_t)k * ny + j) * nx + i;
    if (diag && diag[v] == 0.f) return;
    const float m = s_dev ? s_dev[0] * gain : 1.0f;
    const float invf = 1.0f / (float)f;
    const float gx = fminf(fmaxf((i + 0.5f) * invf - 0.5f, 0.f), (float)(cx - 1));
    const float gy = fminf(fmaxf((j + 0.5f) * invf - 0.5f, 0.f), (float)(cy - 1));
    const float gz = fminf(fmaxf((k + 0.5f) * invf - 0.5f, 0.f), (float)(cz - 1));
    const int i0 = (int)gx, j0 = (int)gy, k0 = (int)gz;
    const int i1 = min(i0 + 1, cx - 1), j1 = min(j0 + 1, cy - 1), k1 = min(k0 + 1, cz - 1);
    const float fx = gx - i0, fy = gy - j0, fz = gz - k0;
    auto at = [&](int a, int b, int e) { return cp[((size_t)e * cy + b) * cx + a]; };
    const float c00 = at(i0, j0, k0) + fx * (at(i1, j0, k0) - at(i0, j0, k0));
    const float c10 = at(i0, j1, k0) + fx * (at(i1, j1, k0) - at(i0, j1, k0));
    const float c01 = at(i0, j0, k1) + fx * (at(i1, j0, k1) - at(i0, j0, k1));
    const float c11 = at(i0, j1, k1) + fx * (at(i1, j1, k1) - at(i0, j1, k1));
    const float c0 = c00 + fy * (c10 - c00), c1 = c01 + fy * (c11 - c01);
    g[v] += (double)((c0 + fz * (c1 - c0)) * m);
}

// k_prolong_add64 for rows of a multiple of 4 cells: a thread takes 4 consecutive
// cells of a row (one 16-byte diag load, two 16-byte guess loads and stores) and
// evaluates each cell's interpolation with the same expression as k_prolong_add64.
__global__ __launch_bounds__(256) void k_prolong_add64_x4(const float* __restrict__ cp, int cx, int cy, int cz,
                                                          int f, const float* __restrict__ s_dev, float gain,
                                                          const float* __restrict__ diag, double* __restrict__ g) {
    const int nx = cx * f, ny = cy * f;
    const int i4 = (blockIdx.x * 256 + threadIdx.x) * 4, j = blockIdx.y, k = blockIdx.z;
    if (i4 >= nx) return;
    const size_t v = ((size_t)k * ny + j) * nx + i4;
    const float m = s_dev ? s_dev[0] * gain : 1.0f;
    const float invf = 1.0f / (float)f;
    const float gy = fminf(fmaxf((j + 0.5f) * invf - 0.5f, 0.f), (float)(cy - 1));
    const float gz = fminf(fmaxf((k + 0.5f) * invf - 0.5f, 0.f), (float)(cz - 1));
    const int j0 = (int)gy, k0 = (int)gz;
    const int j1 = min(j0 + 1, cy - 1), k1 = min(k0 + 1, cz - 1);
    const float fy = gy - j0, fz = gz - k0;
    auto at = [&](int a, int b, int e) { return cp[((size_t)e * cy + b) * cx + a]; };
    float dv[4] = {1.f, 1.f, 1.f, 1.f};
    if (diag) {
        const float4 t = *reinterpret_cast<const float4*>(diag + v);
        dv[0] = t.x; dv[1] = t.y; dv[2] = t.z; dv[3] = t.w;
    }
    const double2 ga = *reinterpret_cast<const double2*>(g + v), gb = *reinterpret_cast<const double2*>(g + v + 2);
    double gv[4] = {ga.x, ga.y, gb.x, gb.y};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float gx = fminf(fmaxf((i4 + c + 0.5f) * invf - 0.5f, 0.f), (float)(cx - 1));
        const int i0 = (int)gx, i1 = min(i0 + 1, cx - 1);
        const float fx = gx - i0;
        const float c00 = at(i0, j0, k0) + fx * (at(i1, j0, k0) - at(i0, j0, k0));
        const float c10 = at(i0, j1, k0) + fx * (at(i1, j1, k0) - at(i0, j1, k0));
        const float c01 = at(i0, j0, k1) + fx * (at(i1, j0, k1) - at(i0, j0, k1));
        const float c11 = at(i0, j1, k1) + fx * (at(i1, j1, k1) - at(i0, j1, k1));
        const float c0 = c00 + fy * (c10 - c00), c1 = c01 + fy * (c11 - c01);
        if (dv[c] != 0.f) gv[c] += (double)((c0 + fz * (c1 - c0)) * m);
    }
    *reinterpret_cast<double2*>(g + v) = make_double2(gv[0], gv[1]);
    *reinterpret_cast<double2*>(g + v + 2) = make_double2(gv[2], gv[3]);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ a, size_t n, bf16* __restrict__ b) {
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) b[e] = (bf16)a[e];
}

}  // namespace unet
}  // namespace vflip
