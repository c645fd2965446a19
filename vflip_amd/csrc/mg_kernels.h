// Device kernels of the multigrid-preconditioned CG solver (included by mgpcg.hip).
//
// Every level is a symmetric 7-point operator: diag + couplings to the +x/+y/+z
// neighbour (ax[c] = A(c, c+1) = A(c+1, c)), linear x-fastest storage, 32-bit
// cell indices (grids up to 2^31 cells).
#pragma once

#include "common.h"

namespace vflip {

// ----------------------------------------------------------------------------
// Operator view. CT = float (FP32 storage) or __half (scaled FP16 storage).
// ----------------------------------------------------------------------------
template <typename CT> struct Op {
    const CT* d;
    const CT* ax;
    const CT* ay;
    const CT* az;
    float scale;
    int ni, nj, nk;
    __device__ inline float D(int c) const { return scale * (float)d[c]; }
    __device__ inline float AX(int c) const { return scale * (float)ax[c]; }
    __device__ inline float AY(int c) const { return scale * (float)ay[c]; }
    __device__ inline float AZ(int c) const { return scale * (float)az[c]; }
};

__device__ inline void unflat32(unsigned c, unsigned ni, unsigned nj, int& i, int& j, int& k) {
    unsigned t = c / ni;
    i = (int)(c - t * ni);
    unsigned kk = t / nj;
    j = (int)(t - kk * nj);
    k = (int)kk;
}

// Off-diagonal row sum in the reference CSR column order (-z, -y, -x, +x, +y, +z),
// S/AlgebraicMultigrid.h:41-53.
template <typename CT>
__device__ inline float offdiag_sum(const Op<CT>& op, const float* __restrict__ x, int c, int i, int j, int k) {
    const int sl = op.ni * op.nj;
    float s = 0.f;
    if (k > 0) s += op.AZ(c - sl) * x[c - sl];
    if (j > 0) s += op.AY(c - op.ni) * x[c - op.ni];
    if (i > 0) s += op.AX(c - 1) * x[c - 1];
    if (i + 1 < op.ni) s += op.AX(c) * x[c + 1];
    if (j + 1 < op.nj) s += op.AY(c) * x[c + op.ni];
    if (k + 1 < op.nk) s += op.AZ(c) * x[c + sl];
    return s;
}

// GS update x = (b - offdiag)/diag, diag == 0 -> 0 (S/AlgebraicMultigrid.h:54-62).
__device__ inline float gs_value(float b, float s, float d) { return (d != 0.f) ? (b - s) / d : 0.f; }

// ----------------------------------------------------------------------------
// Small-level smoother: one red-black half sweep in place (flat launch).
// ----------------------------------------------------------------------------
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_half(Op<CT> op, const float* __restrict__ b, float* __restrict__ x,
                                                      int colour) {
    const unsigned n = (unsigned)(op.ni * op.nj * op.nk);
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat32(c, op.ni, op.nj, i, j, k);
    if (((i + j + k) & 1) != colour) return;
    x[c] = gs_value(b[c], offdiag_sum(op, x, (int)c, i, j, k), op.D(c));
}

// First half sweep from a zero guess: odd cells b/diag, even cells 0.
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_first(Op<CT> op, const float* __restrict__ b, float* __restrict__ x) {
    const unsigned n = (unsigned)(op.ni * op.nj * op.nk);
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat32(c, op.ni, op.nj, i, j, k);
    x[c] = ((i + j + k) & 1) ? gs_value(b[c], 0.f, op.D(c)) : 0.f;
}

// Odd half sweep fused with the preceding prolongation x += P xc: odd cells read
// their even neighbours as x + xc(parent) (S/AlgebraicMultigrid.h:118-129, 161-165).
// Even cells keep the uncorrected value: the next even half sweep overwrites
// them without reading it.
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_odd_prolong(Op<CT> op, const float* __restrict__ b,
                                                             float* __restrict__ x, const float* __restrict__ xc,
                                                             int ci, int cj) {
    const unsigned n = (unsigned)(op.ni * op.nj * op.nk);
    unsigned cu = blockIdx.x * kBlock + threadIdx.x;
    if (cu >= n) return;
    int i, j, k;
    unflat32(cu, op.ni, op.nj, i, j, k);
    if (!((i + j + k) & 1)) return;
    const int c = (int)cu, sl = op.ni * op.nj;
    auto X = [&](int q, int a, int bb, int cc) { return x[q] + xc[((cc >> 1) * cj + (bb >> 1)) * ci + (a >> 1)]; };
    float s = 0.f;
    if (k > 0) s += op.AZ(c - sl) * X(c - sl, i, j, k - 1);
    if (j > 0) s += op.AY(c - op.ni) * X(c - op.ni, i, j - 1, k);
    if (i > 0) s += op.AX(c - 1) * X(c - 1, i - 1, j, k);
    if (i + 1 < op.ni) s += op.AX(c) * X(c + 1, i + 1, j, k);
    if (j + 1 < op.nj) s += op.AY(c) * X(c + op.ni, i, j + 1, k);
    if (k + 1 < op.nk) s += op.AZ(c) * X(c + sl, i, j, k + 1);
    x[c] = gs_value(b[c], s, op.D(c));
}

// x += P xc (injection), used when there are no post-smoothing sweeps.
__global__ __launch_bounds__(kBlock) void k_prolong_add(int ni, int nj, int nk, float* __restrict__ x,
                                                        const float* __restrict__ xc, int ci, int cj) {
    const unsigned n = (unsigned)(ni * nj * nk);
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat32(c, ni, nj, i, j, k);
    x[c] += xc[((k >> 1) * cj + (j >> 1)) * ci + (i >> 1)];
}

// ----------------------------------------------------------------------------
// Large-level smoother: one full red-black sweep per launch, x ping-ponged
// (xin -> xout). A workgroup owns a 64 x TY column tile and a z range and
// marches up it: at plane p it computes the odd (red) cells of p on the tile plus
// a one-cell ring from xin (global, L1/L2 resident), keeps the last three red
// planes in LDS, and then finishes the even (black) cells of plane p-1 from those
// red planes, writing both colours of plane p-1. Ring red cells are recomputed by
// both neighbouring workgroups (bit-identical values), so there is no grid-wide
// barrier between the two half sweeps and each cell's operator, b and x are
// streamed from HBM once per sweep.
//   mode 0: xin is the current iterate
//   mode 1: the current iterate is zero (first pre-smoothing sweep; xin unused)
//   mode 2: the current iterate is xin + P xc (first post-smoothing sweep)
// dot != nullptr accumulates sum(xout * b) per workgroup (z.r in PCG).
// ----------------------------------------------------------------------------
constexpr int kZTX = 64;

template <int TY, typename Acc>
__global__ __launch_bounds__(256) void k_rbgs_zsweep(Op<float> op, const float* __restrict__ b,
                                                     const float* __restrict__ xin, float* __restrict__ xout, int mode,
                                                     const float* __restrict__ xc, int ci, int cj, int tiles_x,
                                                     int tiles_y, int zchunk, double* __restrict__ dot) {
    constexpr int TX = kZTX;
    constexpr int RW = TX + 2, RH = TY + 2;
    constexpr int RSLOTS = (RW / 2) * RH;
    constexpr int BSLOTS = (TX / 2) * TY;
    __shared__ float sr[3][RH * RW];
    const int ni = op.ni, nj = op.nj, nk = op.nk;
    const int sl = ni * nj;
    // spatially adjacent tiles on the same XCD (dispatch round-robins XCDs)
    const int nb = gridDim.x, L = blockIdx.x;
    const int lg = (nb % 8 == 0) ? (L % 8) * (nb / 8) + (L / 8) : L;
    const int tx = lg % tiles_x, t2 = lg / tiles_x;
    const int ty = t2 % tiles_y, zc = t2 / tiles_y;
    const int i0 = tx * TX, j0 = ty * TY, k0 = zc * zchunk;
    const int k1 = min(k0 + zchunk, nk);
    const int tid = threadIdx.x;
    Acc dacc = 0;

    auto xold = [&](int c, int i, int j, int k) -> float {
        float v = xin[c];
        if (mode == 2) v += xc[((k >> 1) * cj + (j >> 1)) * ci + (i >> 1)];
        return v;
    };

    for (int p = k0 - 1; p <= k1; ++p) {
        if (p >= 0 && p < nk) {
            float* R = sr[(p + 3) % 3];
            for (int s = tid; s < RSLOTS; s += 256) {
                const int jr = s / (RW / 2), m = s - jr * (RW / 2);
                const int j = j0 - 1 + jr;
                const int ir = 2 * m + ((i0 + j + p) & 1);
                const int i = i0 - 1 + ir;
                if (i < 0 || i >= ni || j < 0 || j >= nj) continue;
                const int c = (p * nj + j) * ni + i;
                float sum = 0.f;
                if (mode != 1) {
                    if (p > 0) sum += op.AZ(c - sl) * xold(c - sl, i, j, p - 1);
                    if (j > 0) sum += op.AY(c - ni) * xold(c - ni, i, j - 1, p);
                    if (i > 0) sum += op.AX(c - 1) * xold(c - 1, i - 1, j, p);
                    if (i + 1 < ni) sum += op.AX(c) * xold(c + 1, i + 1, j, p);
                    if (j + 1 < nj) sum += op.AY(c) * xold(c + ni, i, j + 1, p);
                    if (p + 1 < nk) sum += op.AZ(c) * xold(c + sl, i, j, p + 1);
                }
                R[jr * RW + ir] = gs_value(b[c], sum, op.D(c));
            }
        }
        __syncthreads();
        const int q = p - 1;
        if (q >= k0 && q < k1) {
            const float* Rm = sr[(q + 2) % 3];
            const float* R0 = sr[(q + 3) % 3];
            const float* Rp = sr[(p + 3) % 3];
            for (int s = tid; s < BSLOTS; s += 256) {
                const int jt = s / (TX / 2), m = s - jt * (TX / 2);
                const int j = j0 + jt;
                if (j >= nj) continue;
                const int it = 2 * m + ((i0 + j + q) & 1);
                const int i = i0 + it;
                const int jr = jt + 1, ir = it + 1;
                const int row = (q * nj + j) * ni;
                if (i < ni) {
                    const int c = row + i;
                    float sum = 0.f;
                    if (q > 0) sum += op.AZ(c - sl) * Rm[jr * RW + ir];
                    if (j > 0) sum += op.AY(c - ni) * R0[(jr - 1) * RW + ir];
                    if (i > 0) sum += op.AX(c - 1) * R0[jr * RW + ir - 1];
                    if (i + 1 < ni) sum += op.AX(c) * R0[jr * RW + ir + 1];
                    if (j + 1 < nj) sum += op.AY(c) * R0[(jr + 1) * RW + ir];
                    if (q + 1 < nk) sum += op.AZ(c) * Rp[jr * RW + ir];
                    const float bc = b[c];
                    const float v = gs_value(bc, sum, op.D(c));
                    xout[c] = v;
                    if (dot) dacc += (Acc)v * (Acc)bc;
                }
                const int ip = i0 + (it ^ 1);
                if (ip < ni) {
                    const int c2 = row + ip;
                    const float vr = R0[jr * RW + (it ^ 1) + 1];
                    xout[c2] = vr;
                    if (dot) dacc += (Acc)vr * (Acc)b[c2];
                }
            }
        }
        __syncthreads();
    }
    if (dot) {
        Acc t = block_sum(dacc);
        if (tid == 0) dot[blockIdx.x] = (double)t;
    }
}

// ----------------------------------------------------------------------------
// b_c = R (b - A x), R = 1/8 per existing child (S/AlgebraicMultigrid.h:105-117,
// S/GeometricLevelGen.cpp:50-63). 3D launch over coarse cells (64 x 4 threads),
// so a wavefront reads 128 consecutive fine cells per row.
// ----------------------------------------------------------------------------
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_residual_restrict(Op<CT> op, const float* __restrict__ b,
                                                              const float* __restrict__ x, float* __restrict__ bc,
                                                              int ci, int cj, int ck) {
    const int I = blockIdx.x * 64 + threadIdx.x, J = blockIdx.y * 4 + threadIdx.y, K = blockIdx.z;
    if (I >= ci || J >= cj || K >= ck) return;
    float acc = 0.f;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                if (i >= op.ni || j >= op.nj || k >= op.nk) continue;
                const int c = (k * op.nj + j) * op.ni + i;
                const float r = b[c] - (offdiag_sum(op, x, c, i, j, k) + op.D(c) * x[c]);
                acc += 0.125f * r;
            }
    bc[(K * cj + J) * ci + I] = acc;
}

// ----------------------------------------------------------------------------
// Coarsest level (<= 4096 unknowns, S/GeometricLevelGen.h:46): all RB-GS sweeps
// in one workgroup, x in LDS, operator in registers. Each thread owns up to 8
// x-pairs (2m, 2m+1) of a row; a pair holds one cell of each colour, so every
// lane has work in both half sweeps.
// ----------------------------------------------------------------------------
constexpr int kCoarseMax = 4096;
constexpr int kCoarseThreads = 256;
constexpr int kCoarsePairs = 8;  // pairs per thread: up to 2048 pairs

template <typename CT>
__global__ __launch_bounds__(kCoarseThreads) void k_coarsest(Op<CT> op, const float* __restrict__ b,
                                                             float* __restrict__ x, int sweeps) {
    __shared__ float xs[kCoarseMax + 1];
    const int ni = op.ni, nj = op.nj, nk = op.nk;
    const int sl = ni * nj;
    const int ppr = (ni + 1) / 2;
    const int npairs = ppr * nj * nk;
    float cd[kCoarsePairs][2], cb[kCoarsePairs][2], ca[kCoarsePairs][2][6];
    int nbr[kCoarsePairs][2][6], cell[kCoarsePairs][2];
#pragma unroll
    for (int q = 0; q < kCoarsePairs; ++q) {
        // empty slots read the zero pad at xs[kCoarseMax] with zero couplings
#pragma unroll
        for (int col = 0; col < 2; ++col) {
            cell[q][col] = -1;
            cd[q][col] = 0.f;
            cb[q][col] = 0.f;
#pragma unroll
            for (int e = 0; e < 6; ++e) {
                nbr[q][col][e] = kCoarseMax;
                ca[q][col][e] = 0.f;
            }
        }
        const int P = threadIdx.x + q * kCoarseThreads;
        if (P >= npairs) continue;
        const int row = P / ppr, m = P - row * ppr;
        const int j = row % nj, k = row / nj;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int i = 2 * m + e;
            if (i >= ni) continue;
            const int c = (k * nj + j) * ni + i;
            const int col = (i + j + k) & 1;
            cell[q][col] = c;
            cd[q][col] = op.D(c);
            cb[q][col] = b[c];
            nbr[q][col][0] = k > 0 ? c - sl : kCoarseMax;
            ca[q][col][0] = k > 0 ? op.AZ(c - sl) : 0.f;
            nbr[q][col][1] = j > 0 ? c - ni : kCoarseMax;
            ca[q][col][1] = j > 0 ? op.AY(c - ni) : 0.f;
            nbr[q][col][2] = i > 0 ? c - 1 : kCoarseMax;
            ca[q][col][2] = i > 0 ? op.AX(c - 1) : 0.f;
            nbr[q][col][3] = i + 1 < ni ? c + 1 : kCoarseMax;
            ca[q][col][3] = i + 1 < ni ? op.AX(c) : 0.f;
            nbr[q][col][4] = j + 1 < nj ? c + ni : kCoarseMax;
            ca[q][col][4] = j + 1 < nj ? op.AY(c) : 0.f;
            nbr[q][col][5] = k + 1 < nk ? c + sl : kCoarseMax;
            ca[q][col][5] = k + 1 < nk ? op.AZ(c) : 0.f;
            xs[c] = 0.f;
        }
    }
    if (threadIdx.x == 0) xs[kCoarseMax] = 0.f;
    __syncthreads();
    for (int it = 0; it < sweeps; ++it) {
#pragma unroll
        for (int col = 1; col >= 0; --col) {
            float nv[kCoarsePairs];
#pragma unroll
            for (int q = 0; q < kCoarsePairs; ++q) {
                float s = 0.f;
#pragma unroll
                for (int e = 0; e < 6; ++e) s += ca[q][col][e] * xs[nbr[q][col][e]];
                nv[q] = s;
            }
#pragma unroll
            for (int q = 0; q < kCoarsePairs; ++q)
                if (cell[q][col] >= 0) xs[cell[q][col]] = gs_value(cb[q][col], nv[q], cd[q][col]);
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < kCoarsePairs; ++q)
#pragma unroll
        for (int col = 0; col < 2; ++col)
            if (cell[q][col] >= 0) x[cell[q][col]] = xs[cell[q][col]];
}

// ----------------------------------------------------------------------------
// Closed-form Galerkin coarsening A_c = 0.5 R A P (R = 1/8, P = injection):
//   coupling(I, I+e) = 1/16 * sum of the fine couplings across the aggregate face
//   leak(I)          = 1/16 * sum of the children's leaks
//   diag(I)          = leak(I) + sum of the 6 coarse couplings' magnitudes
// (equal to 1/16 * sum_{i,k in I} A(i,k); written this way no term cancels).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_coarsen(int ni, int nj, int nk, const float* __restrict__ ax,
                                                    const float* __restrict__ ay, const float* __restrict__ az,
                                                    const float* __restrict__ leak, int ci, int cj, int ck,
                                                    float* __restrict__ dc, float* __restrict__ axc,
                                                    float* __restrict__ ayc, float* __restrict__ azc,
                                                    float* __restrict__ leakc) {
    const unsigned nc = (unsigned)(ci * cj * ck);
    unsigned C = blockIdx.x * kBlock + threadIdx.x;
    if (C >= nc) return;
    int I, J, K;
    unflat32(C, ci, cj, I, J, K);
    const float s = 0.0625f;
    auto face = [&](int axis, int plane) -> float {
        float acc = 0.f;
        for (int u = 0; u < 2; ++u)
            for (int v = 0; v < 2; ++v) {
                int i, j, k;
                if (axis == 0) { i = plane; j = 2 * J + u; k = 2 * K + v; }
                else if (axis == 1) { i = 2 * I + u; j = plane; k = 2 * K + v; }
                else { i = 2 * I + u; j = 2 * J + v; k = plane; }
                if (i < 0 || j < 0 || k < 0 || i >= ni || j >= nj || k >= nk) continue;
                const int c = (k * nj + j) * ni + i;
                acc += (axis == 0 ? ax[c] : axis == 1 ? ay[c] : az[c]);
            }
        return s * acc;
    };
    const float cxp = (I + 1 < ci) ? face(0, 2 * I + 1) : 0.f;
    const float cyp = (J + 1 < cj) ? face(1, 2 * J + 1) : 0.f;
    const float czp = (K + 1 < ck) ? face(2, 2 * K + 1) : 0.f;
    const float cxm = (I > 0) ? face(0, 2 * I - 1) : 0.f;
    const float cym = (J > 0) ? face(1, 2 * J - 1) : 0.f;
    const float czm = (K > 0) ? face(2, 2 * K - 1) : 0.f;
    float lk = 0.f;
    for (int kk = 0; kk < 2; ++kk)
        for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
                const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                if (i < ni && j < nj && k < nk) lk += leak[(k * nj + j) * ni + i];
            }
    lk *= s;
    leakc[C] = lk;
    axc[C] = cxp;
    ayc[C] = cyp;
    azc[C] = czp;
    dc[C] = lk + ((-czm) + (-cym) + (-cxm) + (-cxp) + (-cyp) + (-czp));
}

// leak = diag - sum|couplings| in FP64 (for operators given without it).
__global__ __launch_bounds__(kBlock) void k_leak_from_diag(int ni, int nj, int nk, const float* __restrict__ d,
                                                           const float* __restrict__ ax, const float* __restrict__ ay,
                                                           const float* __restrict__ az, float* __restrict__ leak) {
    const unsigned n = (unsigned)(ni * nj * nk);
    unsigned cu = blockIdx.x * kBlock + threadIdx.x;
    if (cu >= n) return;
    int i, j, k;
    unflat32(cu, ni, nj, i, j, k);
    const int c = (int)cu, sl = ni * nj;
    double s = 0;
    if (k > 0) s -= az[c - sl];
    if (j > 0) s -= ay[c - ni];
    if (i > 0) s -= ax[c - 1];
    if (i + 1 < ni) s -= ax[c];
    if (j + 1 < nj) s -= ay[c];
    if (k + 1 < nk) s -= az[c];
    const double l = (double)d[c] - s;
    leak[c] = (float)(l > 0 ? l : 0.0);
}

// FP16 copy of a level scaled by 1/scale.
__global__ __launch_bounds__(kBlock) void k_to_half(unsigned n, const float* __restrict__ src, __half* __restrict__ dst,
                                                    float inv_scale) {
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) dst[c] = __float2half(src[c] * inv_scale);
}

// ----------------------------------------------------------------------------
// Krylov kernels. Reductions write one partial per workgroup; a single-block
// finalize kernel turns them into the CG scalars on the device. Acc = float
// (plain FP32 mode) or double (mixed mode: FP64 dots and alpha/beta).
// ----------------------------------------------------------------------------
constexpr int kVecBlocks = 2048;
constexpr int kMatvecKZ = 8;

struct Scalars {
    double rho, rho_new, sigma, alpha, beta, res, bnorm, tol_abs, true_res;
    int breakdown;
};

// q = A s and sum(s.q); 3D launch (64 x 4 threads, kMatvecKZ planes per block).
template <typename CT, typename Acc>
__global__ __launch_bounds__(kBlock) void k_matvec_dot(Op<CT> op, const float* __restrict__ s, float* __restrict__ q,
                                                       double* __restrict__ partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int kb = blockIdx.z * kMatvecKZ;
    const int ni = op.ni, nj = op.nj, nk = op.nk, sl = ni * nj;
    Acc acc = 0;
    if (i < ni && j < nj) {
        for (int k = kb; k < min(kb + kMatvecKZ, nk); ++k) {
            const int c = (k * nj + j) * ni + i;
            // reference multiply order: -z, -y, -x, diag, +x, +y, +z
            float y = 0.f;
            if (k > 0) y += op.AZ(c - sl) * s[c - sl];
            if (j > 0) y += op.AY(c - ni) * s[c - ni];
            if (i > 0) y += op.AX(c - 1) * s[c - 1];
            y += op.D(c) * s[c];
            if (i + 1 < ni) y += op.AX(c) * s[c + 1];
            if (j + 1 < nj) y += op.AY(c) * s[c + ni];
            if (k + 1 < nk) y += op.AZ(c) * s[c + sl];
            q[c] = y;
            acc += (Acc)s[c] * (Acc)y;
        }
    }
    const int tflat = threadIdx.y * 64 + threadIdx.x;
    (void)tflat;
    Acc t = block_sum(acc);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = (double)t;
}

// r = b - A x in FP64 (true residual); partial max|r|. r may be null.
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_true_residual(Op<CT> op, const float* __restrict__ b,
                                                          const float* __restrict__ x, float* __restrict__ r,
                                                          double* __restrict__ partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int kb = blockIdx.z * kMatvecKZ;
    const int ni = op.ni, nj = op.nj, nk = op.nk, sl = ni * nj;
    double m = 0;
    if (i < ni && j < nj) {
        for (int k = kb; k < min(kb + kMatvecKZ, nk); ++k) {
            const int c = (k * nj + j) * ni + i;
            double y = (double)op.D(c) * x[c];
            if (k > 0) y += (double)op.AZ(c - sl) * x[c - sl];
            if (j > 0) y += (double)op.AY(c - ni) * x[c - ni];
            if (i > 0) y += (double)op.AX(c - 1) * x[c - 1];
            if (i + 1 < ni) y += (double)op.AX(c) * x[c + 1];
            if (j + 1 < nj) y += (double)op.AY(c) * x[c + ni];
            if (k + 1 < nk) y += (double)op.AZ(c) * x[c + sl];
            const double rv = (double)b[c] - y;
            if (r) r[c] = (float)rv;
            m = fmax(m, fabs(rv));
        }
    }
    double t = block_max(m);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = t;
}

template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_dot(unsigned n, const float* __restrict__ a, const float* __restrict__ b,
                                                double* __restrict__ partial) {
    Acc acc = 0;
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) acc += (Acc)a[c] * (Acc)b[c];
    Acc t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

__global__ __launch_bounds__(kBlock) void k_absmax(unsigned n, const float* __restrict__ a, double* __restrict__ partial) {
    float m = 0.f;
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) m = fmaxf(m, fabsf(a[c]));
    float t = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

// x += alpha s; r -= alpha q; partial max|r| (S/AlgebraicMultigrid.h:259-262)
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_update_xr(unsigned n, const Scalars* __restrict__ sc, float* __restrict__ x,
                                                      const float* __restrict__ s, float* __restrict__ r,
                                                      const float* __restrict__ q, double* __restrict__ partial) {
    const Acc alpha = (Acc)sc->alpha;
    float m = 0.f;
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) {
        x[c] = (float)((Acc)x[c] + alpha * (Acc)s[c]);
        const float rv = (float)((Acc)r[c] - alpha * (Acc)q[c]);
        r[c] = rv;
        m = fmaxf(m, fabsf(rv));
    }
    float t = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

// s = z + beta s (S/AlgebraicMultigrid.h:283)
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_update_s(unsigned n, const Scalars* __restrict__ sc, float* __restrict__ s,
                                                     const float* __restrict__ z) {
    const Acc beta = (Acc)sc->beta;
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock)
        s[c] = (float)((Acc)z[c] + beta * (Acc)s[c]);
}

template <typename Acc> __device__ inline double reduce_partials_sum(const double* p, int np) {
    Acc a = 0;
    for (int q = threadIdx.x; q < np; q += kBlock) a += (Acc)p[q];
    return (double)block_sum(a);
}
__device__ inline double reduce_partials_max(const double* p, int np) {
    double a = 0;
    for (int q = threadIdx.x; q < np; q += kBlock) a = fmax(a, p[q]);
    return block_max(a);
}

// what: 0 rho; 1 sigma -> alpha; 2 res; 3 rho_new -> beta; 4 |b| and tolerance; 5 max into true_res
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_finalize(Scalars* sc, const double* __restrict__ partial, int np, int what,
                                                     double tol) {
    const double v = (what == 2 || what == 4 || what == 5) ? reduce_partials_max(partial, np)
                                                           : reduce_partials_sum<Acc>(partial, np);
    if (threadIdx.x != 0) return;
    if (what == 0) {
        sc->rho = v;
        sc->breakdown = (v == 0 || v != v) ? 1 : 0;
    } else if (what == 1) {
        sc->sigma = v;
        sc->alpha = (double)((Acc)sc->rho / (Acc)v);
    } else if (what == 2) {
        sc->res = v;
    } else if (what == 3) {
        sc->rho_new = v;
        sc->beta = (double)((Acc)v / (Acc)sc->rho);
        sc->rho = v;
    } else if (what == 4) {
        sc->bnorm = v;
        sc->tol_abs = tol * v;
        sc->res = v;
    } else {
        sc->true_res = v;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill(unsigned n, float* __restrict__ a, float v) {
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) a[c] = v;
}
__global__ __launch_bounds__(kBlock) void k_copy(unsigned n, float* __restrict__ dst, const float* __restrict__ src) {
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) dst[c] = src[c];
}

}  // namespace vflip
