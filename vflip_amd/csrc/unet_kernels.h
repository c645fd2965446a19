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

constexpr int kT = 8;            // output tile edge (voxels)
constexpr int kH = kT + 2;       // halo tile edge
constexpr int kHV = kH * kH * kH;
constexpr int kConvThreads = 512;  // 8 waves, wave w owns z-slice w of the tile

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
};

__device__ inline int tap_offset_hv(int tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    return (dz * kH + dy) * kH + dx;
}

// Stages the halo tile of input channels [c0, c0 + CK) into LDS as
// [halo voxel][CK + 8] bf16 (the 8-element pad keeps 16-lane row reads free of
// bank conflicts). Voxels outside the grid read as zero (zero padding).
template <int CK>
__device__ inline void stage_halo(const ConvArgs& a, int c0, int x0, int y0, int z0, bf16* lds) {
    int s = 0, cl = c0;
    if (a.nsrc > 1 && c0 >= a.src[0].C) {
        s = 1;
        cl = c0 - a.src[0].C;
    }
    const ConvSrc src = a.src[s];
    const int sx = a.nx >> src.up, sy = a.ny >> src.up;
    constexpr int P = CK / 8;  // 16-byte parts per voxel
    for (int e = threadIdx.x; e < kHV * P; e += kConvThreads) {
        const int hv = e / P, part = e % P;
        const int hx = hv % kH, hy = (hv / kH) % kH, hz = hv / (kH * kH);
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        bf16x8 val;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < a.nx && gy < a.ny && gz < a.nz) {
            const size_t v = ((size_t)(gz >> src.up) * sy + (gy >> src.up)) * sx + (gx >> src.up);
            val = *reinterpret_cast<const bf16x8*>(src.x + v * src.C + cl + part * 8);
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) val[q] = (bf16)0.f;
        }
        *reinterpret_cast<bf16x8*>(lds + hv * (CK + 8) + part * 8) = val;
    }
}

// 3x3x3 convolution, implicit GEMM: M = 512 voxels of an 8^3 tile, N = cout,
// K = 27 * cin. Each K step of 32 is one tap x 32 channels (CK = 32) or two
// taps x 16 channels (CK = 16). A workgroup computes NB blocks of 16 output
// channels, group (blockIdx / ntiles) % groups of the a.groups groups (cout =
// 16 NB groups); small grids split their channels over groups so that more
// workgroups share a layer. With ksplit > 1 workgroup slice blockIdx /
// (ntiles groups) takes a contiguous range of the K steps (channel chunk x
// step, the unsplit order) and writes its FP32 partial sums instead of the
// outputs; with ksplit 1 the loop is the unsplit one.
template <int CK, int NB>
__global__ __launch_bounds__(kConvThreads) void k_conv3(ConvArgs a) {
    __shared__ __attribute__((aligned(16))) bf16 lds[kHV * (CK + 8)];
    const int L = blockIdx.x % a.ntiles, grp = (blockIdx.x / a.ntiles) % a.groups;
    const int ks = blockIdx.x / (a.ntiles * a.groups), nks = a.ksplit > 1 ? a.ksplit : 1;
    const int cbase = grp * NB * 16;
    int t = L;
    if (a.ntiles % 8 == 0) t = (L % 8) * (a.ntiles / 8) + L / 8;
    const int x0 = (t % a.tiles_x) * kT, y0 = ((t / a.tiles_x) % a.tiles_y) * kT, z0 = (t / (a.tiles_x * a.tiles_y)) * kT;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kg = lane >> 4;  // MFMA row/col within a block, k group of 8
    f32x4 acc[4][NB];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // halo voxel of this lane's output row in each M block (tap 0 offset added per tap)
    int hv0[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const int m = mb * 16 + r16;
        hv0[mb] = ((w + 1) * kH + (m / 8 + 1)) * kH + (m % 8 + 1);
    }
    constexpr int STEPS = CK == 32 ? 27 : 14;
    const int total = ((a.cin + CK - 1) / CK) * STEPS;
    const int q0 = (int)((long long)ks * total / nks), q1 = (int)((long long)(ks + 1) * total / nks);
    // K step q: tap and channel offset of this lane's 8 k values
    auto step_k = [&](int q, int& tap, int& coff) {
        const int st = q % STEPS;
        if (CK == 32) {
            tap = st;
            coff = kg * 8;
        } else {
            tap = 2 * st + (kg >> 1);  // tap 27 is the zero pad of the weights
            coff = (kg & 1) * 8;
        }
    };
    // the weight fragments of step q (global loads; issued one step ahead so that
    // their latency overlaps the previous step's matrix-core work)
    auto load_b = [&](int q, bf16x8 (&bf)[NB]) {
        int tap, coff;
        step_k(q, tap, coff);
        const size_t kidx = (size_t)tap * a.cin + (q / STEPS) * CK + coff;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
            bf[nb] = *reinterpret_cast<const bf16x8*>(a.w + (size_t)(cbase + nb * 16 + r16) * a.kw + kidx);
    };
    bf16x8 bcur[NB], bnext[NB];
    if (q0 < q1) load_b(q0, bcur);
    int staged = -1;
    for (int q = q0; q < q1; ++q) {
        const int c0 = (q / STEPS) * CK;
        if (q + 1 < q1) load_b(q + 1, bnext);
        if (c0 != staged) {  // uniform over the workgroup
            __syncthreads();
            stage_halo<CK>(a, c0, x0, y0, z0, lds);
            __syncthreads();
            staged = c0;
        }
        int tap, coff;
        step_k(q, tap, coff);
        const int dhv = tap_offset_hv(tap < 27 ? tap : 0);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const bf16x8 afrag = *reinterpret_cast<const bf16x8*>(lds + (hv0[mb] + dhv) * (CK + 8) + coff);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bcur[nb], acc[mb][nb], 0, 0, 0);
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) bcur[nb] = bnext[nb];
    }
    // epilogue: D reg r of lane -> voxel row kg*4 + r of the M block, channel nb*16 + r16
    const int z = z0 + w;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = mb * 16 + kg * 4 + r;
            const int x = x0 + m % 8, y = y0 + m / 8;
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
                if (a.bias) val += a.bias[co];
                if (a.relu) val = val > 0.f ? val : 0.f;
                if (a.out) a.out[v * a.cout + co] = (bf16)val;
                if (a.out32) {
                    if (!a.out32_ch0) a.out32[v * a.cout + co] = val;
                    else if (co == 0) a.out32[v] = val;
                }
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

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ a, size_t n, bf16* __restrict__ b) {
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) b[e] = (bf16)a[e];
}

}  // namespace unet
}  // namespace vflip
