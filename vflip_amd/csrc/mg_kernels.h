// Device kernels of the multigrid-preconditioned CG solver (included by mgpcg.hip).
//
// Every level is a symmetric 7-point operator: diag + couplings to the +x/+y/+z
// neighbour (ax[c] = A(c, c+1) = A(c+1, c)), linear x-fastest storage, 32-bit
// cell indices (grids up to 2^31 cells).
#pragma once

#include <type_traits>

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
    if (k > 0) s = __builtin_fmaf(op.AZ(c - sl), x[c - sl], s);
    if (j > 0) s = __builtin_fmaf(op.AY(c - op.ni), x[c - op.ni], s);
    if (i > 0) s = __builtin_fmaf(op.AX(c - 1), x[c - 1], s);
    if (i + 1 < op.ni) s = __builtin_fmaf(op.AX(c), x[c + 1], s);
    if (j + 1 < op.nj) s = __builtin_fmaf(op.AY(c), x[c + op.ni], s);
    if (k + 1 < op.nk) s = __builtin_fmaf(op.AZ(c), x[c + sl], s);
    return s;
}

// GS update x = (b - offdiag)/diag, diag == 0 -> 0 (S/AlgebraicMultigrid.h:54-62).
__device__ inline float gs_value(float b, float s, float d) { return (d != 0.f) ? (b - s) / d : 0.f; }

// Block coordinates and grid size. Kernel bodies take them explicitly so that a
// batched launch (several subdomains' grids in one launch) can run a body with
// each subdomain's own coordinates; the single-grid kernels pass blockIdx/gridDim.
struct Blk {
    unsigned x, y, z, nx, ny, nz;
};
__device__ inline Blk this_blk() { return Blk{blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z}; }

// ----------------------------------------------------------------------------
// Small-level smoother: one red-black half sweep in place (flat launch).
// ----------------------------------------------------------------------------
template <typename CT>
__device__ inline void rbgs_half_body(const Op<CT>& op, const float* __restrict__ b, float* __restrict__ x, int colour,
                                      Blk B) {
    const unsigned n = (unsigned)(op.ni * op.nj * op.nk);
    unsigned c = B.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat32(c, op.ni, op.nj, i, j, k);
    if (((i + j + k) & 1) != colour) return;
    x[c] = gs_value(b[c], offdiag_sum(op, x, (int)c, i, j, k), op.D(c));
}
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_half(Op<CT> op, const float* __restrict__ b, float* __restrict__ x,
                                                      int colour) {
    rbgs_half_body(op, b, x, colour, this_blk());
}

// First half sweep from a zero guess: odd cells b/diag, even cells 0.
template <typename CT>
__device__ inline void rbgs_first_body(const Op<CT>& op, const float* __restrict__ b, float* __restrict__ x, Blk B) {
    const unsigned n = (unsigned)(op.ni * op.nj * op.nk);
    unsigned c = B.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    unflat32(c, op.ni, op.nj, i, j, k);
    x[c] = ((i + j + k) & 1) ? gs_value(b[c], 0.f, op.D(c)) : 0.f;
}
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_first(Op<CT> op, const float* __restrict__ b, float* __restrict__ x) {
    rbgs_first_body(op, b, x, this_blk());
}

// Odd half sweep fused with the preceding prolongation x += P xc: odd cells read
// their even neighbours as x + xc(parent) (S/AlgebraicMultigrid.h:118-129, 161-165).
// Even cells keep the uncorrected value: the next even half sweep overwrites
// them without reading it.
template <typename CT>
__device__ inline void rbgs_odd_prolong_body(const Op<CT>& op, const float* __restrict__ b, float* __restrict__ x,
                                             const float* __restrict__ xc, int ci, int cj, Blk B) {
    const unsigned n = (unsigned)(op.ni * op.nj * op.nk);
    unsigned cu = B.x * kBlock + threadIdx.x;
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
template <typename CT>
__global__ __launch_bounds__(kBlock) void k_rbgs_odd_prolong(Op<CT> op, const float* __restrict__ b,
                                                             float* __restrict__ x, const float* __restrict__ xc,
                                                             int ci, int cj) {
    rbgs_odd_prolong_body(op, b, x, xc, ci, cj, this_blk());
}

// x += P xc (injection), used when there are no post-smoothing sweeps.
static __global__ __launch_bounds__(kBlock) void k_prolong_add(int ni, int nj, int nk, float* __restrict__ x,
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
// a one-cell ring from xin, keeps the last three red planes in LDS, and then
// finishes the even (black) cells of plane p-1 from those red planes, writing
// both colours of plane p-1. Ring red cells are recomputed by both neighbouring
// workgroups (bit-identical values), so there is no grid-wide barrier between the
// two half sweeps and each cell's operator, b and x are streamed from HBM once
// per sweep. The loop is software-pipelined: the global inputs of the next
// plane's red cells and of this plane's black cells are loaded one step ahead,
// so each step's memory latency overlaps the previous step's work.
//   MODE 0: xin is the current iterate
//   MODE 1: the current iterate is zero (first pre-smoothing sweep; xin unused)
//   MODE 2: the current iterate is xin + P xc (first post-smoothing sweep)
// dot != nullptr accumulates sum(xout * b) per workgroup (z.r in PCG).
// ----------------------------------------------------------------------------
constexpr int kZTX = 64;

struct RedIn {
    float b, d, a[6], x[6];
    int slot;  // LDS index, -1 when the slot has no cell
};
struct BlackIn {
    float b, d, a[6], bp;
    int c;  // cell index, -1 when absent
    int c2; // partner (odd) cell index, -1 when absent
};

template <int TY, int MODE, typename Acc, typename CT>
__device__ inline void rbgs_zsweep_body(const Op<CT>& op, const float* __restrict__ b,
                                        const float* __restrict__ xin, float* __restrict__ xout,
                                        const float* __restrict__ xc, int ci, int cj, int tiles_x, int tiles_y,
                                        int zchunk, double* __restrict__ dot, Blk B) {
    constexpr int TX = kZTX;
    constexpr int RW = TX + 2, RH = TY + 2;
    constexpr int RHALF = RW / 2;
    constexpr int RSLOTS = RHALF * RH;
    constexpr int RPER = (RSLOTS + 255) / 256;
    constexpr int BSLOTS = (TX / 2) * TY;
    constexpr int BPER = (BSLOTS + 255) / 256;
    __shared__ float sr[3][RH * RW];
    const int ni = op.ni, nj = op.nj, nk = op.nk;
    const int sl = ni * nj;
    // spatially adjacent tiles on the same XCD (dispatch round-robins XCDs)
    const int nb = B.nx, L = B.x;
    const int lg = (nb % 8 == 0) ? (L % 8) * (nb / 8) + (L / 8) : L;
    const int tx = lg % tiles_x, t2 = lg / tiles_x;
    const int ty = t2 % tiles_y, zc = t2 / tiles_y;
    const int i0 = tx * TX, j0 = ty * TY, k0 = zc * zchunk;
    const int k1 = min(k0 + zchunk, nk);
    const int tid = threadIdx.x;
    Acc dacc = 0;

    auto xold = [&](int c, int i, int j, int k) -> float {
        float v = xin[c];
        if (MODE == 2) v += xc[((k >> 1) * cj + (j >> 1)) * ci + (i >> 1)];
        return v;
    };
    // Loads are branch-free (clamped in-bounds addresses, masked values) so the
    // compiler can keep the prefetched loads of the next step in flight.
    auto load_red = [&](int p, int s, RedIn& in) {
        const int sc = s < RSLOTS ? s : 0;
        const int jr = sc / RHALF, m = sc - jr * RHALF;
        const int j = j0 - 1 + jr;
        const int ir = 2 * m + ((i0 + j + p) & 1);
        const int i = i0 - 1 + ir;
        const bool ok = s < RSLOTS && p >= 0 && p < nk && i >= 0 && i < ni && j >= 0 && j < nj;
        const int c = ok ? (p * nj + j) * ni + i : 0;
        const int pi = ok ? i : 0, pj = ok ? j : 0, pp = ok ? p : 0;
        in.slot = ok ? jr * RW + ir : -1;
        const bool zm = pp > 0, ym = pj > 0, xm = pi > 0, xp = pi + 1 < ni, yp = pj + 1 < nj, zp = pp + 1 < nk;
        const int czm = zm ? c - sl : c, cym = ym ? c - ni : c, cxm = xm ? c - 1 : c;
        const int cxp = xp ? c + 1 : c, cyp = yp ? c + ni : c, czp = zp ? c + sl : c;
        in.b = b[c];
        in.d = op.D(c);
        const float a0 = op.AZ(czm), a1 = op.AY(cym), a2 = op.AX(cxm), a3 = op.AX(c), a4 = op.AY(c), a5 = op.AZ(c);
        in.a[0] = zm ? a0 : 0.f;
        in.a[1] = ym ? a1 : 0.f;
        in.a[2] = xm ? a2 : 0.f;
        in.a[3] = xp ? a3 : 0.f;
        in.a[4] = yp ? a4 : 0.f;
        in.a[5] = zp ? a5 : 0.f;
        if (MODE != 1) {
            float x0 = xin[czm], x1 = xin[cym], x2 = xin[cxm], x3 = xin[cxp], x4 = xin[cyp], x5 = xin[czp];
            if (MODE == 2) {
                auto par = [&](int ii, int jj, int kk) { return ((kk >> 1) * cj + (jj >> 1)) * ci + (ii >> 1); };
                x0 += xc[par(pi, pj, zm ? pp - 1 : pp)];
                x1 += xc[par(pi, ym ? pj - 1 : pj, pp)];
                x2 += xc[par(xm ? pi - 1 : pi, pj, pp)];
                x3 += xc[par(xp ? pi + 1 : pi, pj, pp)];
                x4 += xc[par(pi, yp ? pj + 1 : pj, pp)];
                x5 += xc[par(pi, pj, zp ? pp + 1 : pp)];
            }
            in.x[0] = zm ? x0 : 0.f;
            in.x[1] = ym ? x1 : 0.f;
            in.x[2] = xm ? x2 : 0.f;
            in.x[3] = xp ? x3 : 0.f;
            in.x[4] = yp ? x4 : 0.f;
            in.x[5] = zp ? x5 : 0.f;
        }
    };
    auto load_black = [&](int q, int s, BlackIn& in) {
        const int sc = s < BSLOTS ? s : 0;
        const int jt = sc / (TX / 2), m = sc - jt * (TX / 2);
        const int j = j0 + jt;
        const int it = 2 * m + ((i0 + j + q) & 1);
        const int i = i0 + it, ip = i0 + (it ^ 1);
        const bool rowok = s < BSLOTS && q >= k0 && q < k1 && j < nj;
        const bool ok = rowok && i < ni, okp = rowok && ip < ni;
        const int row = rowok ? (q * nj + j) * ni : 0;
        const int c = ok ? row + i : 0;
        in.c = ok ? c : -1;
        in.c2 = okp ? row + ip : -1;
        in.bp = dot ? b[okp ? row + ip : 0] : 0.f;
        const int pi = ok ? i : 0, pj = ok ? j : 0, pq = ok ? q : 0;
        const bool zm = pq > 0, ym = pj > 0, xm = pi > 0, xp = pi + 1 < ni, yp = pj + 1 < nj, zp = pq + 1 < nk;
        in.b = b[c];
        in.d = op.D(c);
        const float a0 = op.AZ(zm ? c - sl : c), a1 = op.AY(ym ? c - ni : c), a2 = op.AX(xm ? c - 1 : c);
        const float a3 = op.AX(c), a4 = op.AY(c), a5 = op.AZ(c);
        in.a[0] = zm ? a0 : 0.f;
        in.a[1] = ym ? a1 : 0.f;
        in.a[2] = xm ? a2 : 0.f;
        in.a[3] = xp ? a3 : 0.f;
        in.a[4] = yp ? a4 : 0.f;
        in.a[5] = zp ? a5 : 0.f;
    };

    for (int s = tid; s < 3 * RH * RW; s += 256) (&sr[0][0])[s] = 0.f;
    lds_barrier();
    RedIn rin[RPER], rnext[RPER];
    BlackIn bin[BPER], bnext[BPER];
#pragma unroll
    for (int r = 0; r < RPER; ++r) load_red(k0 - 1, tid + 256 * r, rin[r]);
#pragma unroll
    for (int r = 0; r < BPER; ++r) bin[r].c = bin[r].c2 = -1;

    for (int p = k0 - 1; p <= k1; ++p) {
        // prefetch: red inputs of plane p+1, black inputs of plane p
#pragma unroll
        for (int r = 0; r < RPER; ++r) load_red(p + 1 <= k1 ? p + 1 : nk, tid + 256 * r, rnext[r]);
#pragma unroll
        for (int r = 0; r < BPER; ++r) load_black(p, tid + 256 * r, bnext[r]);
        // red cells of plane p
        float* R = sr[(p + 3) % 3];
#pragma unroll
        for (int r = 0; r < RPER; ++r) {
            if (rin[r].slot < 0) continue;
            float s = 0.f;
            if (MODE != 1) {
#pragma unroll
                for (int e = 0; e < 6; ++e) s = __builtin_fmaf(rin[r].a[e], rin[r].x[e], s);
            }
            R[rin[r].slot] = gs_value(rin[r].b, s, rin[r].d);
        }
        lds_barrier();
        // black cells of plane q = p - 1 (inputs loaded during the previous step)
        const int q = p - 1;
        if (q >= k0 && q < k1) {
            const float* Rm = sr[(q + 2) % 3];
            const float* R0 = sr[(q + 3) % 3];
            const float* Rp = sr[(p + 3) % 3];
#pragma unroll
            for (int r = 0; r < BPER; ++r) {
                const int s = tid + 256 * r;
                if (s >= BSLOTS) continue;
                const int jt = s / (TX / 2), m = s - jt * (TX / 2);
                const int j = j0 + jt;
                const int it = 2 * m + ((i0 + j + q) & 1);
                const int jr = jt + 1, ir = it + 1;
                if (bin[r].c >= 0) {
                    // red neighbours outside the grid were never computed: skip them
                    const int i = i0 + it;
                    float sum = 0.f;
                    if (q > 0) sum = __builtin_fmaf(bin[r].a[0], Rm[jr * RW + ir], sum);
                    if (j > 0) sum = __builtin_fmaf(bin[r].a[1], R0[(jr - 1) * RW + ir], sum);
                    if (i > 0) sum = __builtin_fmaf(bin[r].a[2], R0[jr * RW + ir - 1], sum);
                    if (i + 1 < ni) sum = __builtin_fmaf(bin[r].a[3], R0[jr * RW + ir + 1], sum);
                    if (j + 1 < nj) sum = __builtin_fmaf(bin[r].a[4], R0[(jr + 1) * RW + ir], sum);
                    if (q + 1 < nk) sum = __builtin_fmaf(bin[r].a[5], Rp[jr * RW + ir], sum);
                    const float v = gs_value(bin[r].b, sum, bin[r].d);
                    xout[bin[r].c] = v;
                    if (dot) dacc += (Acc)v * (Acc)bin[r].b;
                }
                if (bin[r].c2 >= 0) {
                    const float vr = R0[jr * RW + (it ^ 1) + 1];
                    xout[bin[r].c2] = vr;
                    if (dot) dacc += (Acc)vr * (Acc)bin[r].bp;
                }
            }
        }
        lds_barrier();
#pragma unroll
        for (int r = 0; r < RPER; ++r) rin[r] = rnext[r];
#pragma unroll
        for (int r = 0; r < BPER; ++r) bin[r] = bnext[r];
    }
    if (dot) {
        Acc t = block_sum(dacc);
        if (tid == 0) dot[L] = (double)t;
    }
}
template <int TY, int MODE, typename Acc, typename CT>
__global__ __launch_bounds__(256) void k_rbgs_zsweep(Op<CT> op, const float* __restrict__ b,
                                                     const float* __restrict__ xin, float* __restrict__ xout,
                                                     const float* __restrict__ xc, int ci, int cj, int tiles_x,
                                                     int tiles_y, int zchunk, double* __restrict__ dot) {
    rbgs_zsweep_body<TY, MODE, Acc, CT>(op, b, xin, xout, xc, ci, cj, tiles_x, tiles_y, zchunk, dot, this_blk());
}

// ----------------------------------------------------------------------------
// Row-wave form of the fused sweep, for levels whose rows are 64*V cells
// (V = 2 or 4). A wave owns one whole x row: lane l holds cells V*l .. V*l+V-1
// and reads each field of the row with one vector load. A 1024-thread workgroup
// stacks kZRows rows: kZOut output rows and one ring row on each side whose red
// cells are recomputed (as in k_rbgs_zsweep), and marches up a z chunk. At plane
// p every wave computes the red cells of its row, then the black cells of plane
// p-1. Within a row the colours alternate, so with e = (j + p) & 1 the red cells
// of plane p and the black cells of plane p-1 both sit at c = 2h + 1 - e
// (h < V/2) and the black cells of plane p at c = 2h + e; every pattern below is
// a select on e, uniform per wave. Neighbour values across lanes come from
// shuffles, across rows from LDS (double-buffered by plane parity: one barrier
// per plane), across planes from registers. Fields are loaded one plane ahead and
// x two planes ahead. Both kernels accumulate the same terms in the same order
// with explicit fused multiply-adds, so they produce the same iterate bit for bit.
// ----------------------------------------------------------------------------
constexpr int kZRows = 16;
constexpr int kZOut = kZRows - 2;

template <int N> __device__ inline void vld(const float* p, float (&v)[N]) {
    if constexpr (N == 1) {
        v[0] = p[0];
    } else if constexpr (N == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p);
        v[0] = t.x; v[1] = t.y;
    } else {
        const float4 t = *reinterpret_cast<const float4*>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    }
}
// FP16 coefficients (the scaled copies of a mixed-precision hierarchy), widened
// to float; the caller applies the level's scale.
template <int N> __device__ inline void vld(const __half* p, float (&v)[N]) {
    if constexpr (N == 1) {
        v[0] = __half2float(p[0]);
    } else if constexpr (N == 2) {
        const float2 t = __half22float2(*reinterpret_cast<const __half2*>(p));
        v[0] = t.x; v[1] = t.y;
    } else {
        const uint2 u = *reinterpret_cast<const uint2*>(p);
        __half2 h0, h1;
        __builtin_memcpy(&h0, &u.x, 4);
        __builtin_memcpy(&h1, &u.y, 4);
        const float2 t0 = __half22float2(h0), t1 = __half22float2(h1);
        v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y;
    }
}
template <int N> __device__ inline void vst(float* p, const float (&v)[N]) {
    if constexpr (N == 1) {
        p[0] = v[0];
    } else if constexpr (N == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// Components of a row vector at c = 2h + 1 - E (Rc) and c = 2h + E (Bc).
template <int E, int V> __device__ inline float Rc(const float (&v)[V], int h) { return v[2 * h + 1 - E]; }
template <int E, int V> __device__ inline float Bc(const float (&v)[V], int h) { return v[2 * h + E]; }

template <int V, int MODE, typename Acc, typename CT = float>
__global__ __launch_bounds__(1024) void k_rbgs_zrow(Op<CT> op, const float* __restrict__ b,
                                                    const float* __restrict__ xin, float* __restrict__ xout,
                                                    const float* __restrict__ xc, int ci, int cj, int tiles_y,
                                                    int zchunk, double* __restrict__ dot) {
    constexpr int H = V / 2;
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    __shared__ __align__(16) float xs[2][kZRows][64 * H];   // black x values of plane p, per row
    __shared__ __align__(16) float rs[2][kZRows][64 * H];   // red values of plane p, per row
    __shared__ __align__(16) float ays[2][kZRows][64 * V];  // ay of plane p, per row
    __shared__ Acc part[kZRows];
    const int ni = op.ni, nj = op.nj, nk = op.nk;
    const float sc = op.scale;
    // consecutive tiles (the y-neighbours of one z chunk) on one XCD: block L runs on XCD L % 8
    const int nb = gridDim.x, L = blockIdx.x, xcd = L % 8;
    const int lg = xcd * (nb / 8) + min(xcd, nb % 8) + L / 8;
    const int ty = lg % tiles_y, zch = lg / tiles_y;
    const int j0 = ty * kZOut, k0 = zch * zchunk, k1 = min(k0 + zchunk, nk);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int j = j0 - 1 + w;
    const bool row_ok = j >= 0 && j < nj;
    const bool out_row = row_ok && w >= 1 && w <= kZOut;
    const bool ym = j > 0, yp = j + 1 < nj;
    const int jc = min(max(j, 0), nj - 1);
    // ring rows read one row further out: j0-2 below the first, j0+kZOut+1 above the last
    const bool has_ex = w == 0 || w == kZRows - 1;
    const int jx = min(max(w == 0 ? j - 1 : j + 1, 0), nj - 1);
    const int i0 = lane * V;
    const int lprev = (lane + 63) & 63, lnext = (lane + 1) & 63;
    Acc dacc = 0;

    auto at = [&](int jj, int k) { return (min(max(k, 0), nk - 1) * nj + jj) * ni + i0; };
    auto load_x = [&](int jj, int k, float (&v)[V]) {
        const int kk = min(max(k, 0), nk - 1);
        vld<V>(xin + (kk * nj + jj) * ni + i0, v);
        if constexpr (MODE == 2) {
            float cv[H];
            vld<H>(xc + ((kk >> 1) * cj + (jj >> 1)) * ci + (i0 >> 1), cv);
#pragma unroll
            for (int c = 0; c < V; ++c) v[c] += cv[c >> 1];
        }
    };

    float cb[V], cd[V], cx[V], cy[V], cz[V];  // fields of plane p
    float zprev[V];                           // az of plane p-1
    float xbm[H], xb0[H], xbp[H];             // black x values of planes p-1, p, p+1
    float exx[V], exy[V];                     // outer row's x and ay at plane p (ring waves)
    float qb[H], qd[H], qxl[H], qxr[H], qy[H], qyd[H], qz[H], qzd[H], qbr[H];  // black cells of plane p-1
    float nqm[H], nq[H];                      // red values of planes p-2, p-1
    {
        const int k = k0 - 1, o = at(jc, k);
        vld<V>(b + o, cb); vld<V>(op.d + o, cd); vld<V>(op.ax + o, cx); vld<V>(op.ay + o, cy);
        vld<V>(op.az + o, cz);
        vld<V>(op.az + at(jc, k - 1), zprev);
        if constexpr (MODE != 1) {
            // black x values of planes k-1, k, k+1 (at c = 2h + ((j + plane) & 1))
            float t0[V], t1[V], t2[V];
            load_x(jc, k - 1, t0);
            load_x(jc, k, t1);
            load_x(jc, k + 1, t2);
            auto pick = [&](auto E) {
                constexpr int e = decltype(E)::value;  // parity of plane k
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    xbm[h] = Bc<1 - e>(t0, h);
                    xb0[h] = Bc<e>(t1, h);
                    xbp[h] = Bc<1 - e>(t2, h);
                }
            };
            if ((j + k) & 1) pick(I1()); else pick(I0());
            if (has_ex) load_x(jx, k, exx);
        }
        if (has_ex) vld<V>(op.ay + at(jx, k), exy);
#pragma unroll
        for (int h = 0; h < H; ++h) nq[h] = nqm[h] = 0.f;
    }

    for (int p = k0 - 1; p <= k1; ++p) {
        const int buf = p & 1;
        // next plane's fields, x two planes ahead, the ring rows' outer row
        float nb_[V], nd[V], nx[V], ny[V], nz[V], nxx[V], nex[V], ney[V];
        {
            // the last plane reloads planes already read (nothing beyond k1+1 is needed)
            const int pn = p < k1 ? p + 1 : p, px = p < k1 ? p + 2 : p + 1;
            const int o = at(jc, pn);
            vld<V>(b + o, nb_); vld<V>(op.d + o, nd); vld<V>(op.ax + o, nx); vld<V>(op.ay + o, ny);
            vld<V>(op.az + o, nz);
            if constexpr (MODE != 1) {
                load_x(jc, px, nxx);
                if (has_ex) load_x(jx, pn, nex);
            }
            if (has_ex) vld<V>(op.ay + at(jx, pn), ney);
        }
        if constexpr (MODE != 1) vst<H>(&xs[buf][w][lane * H], xb0);
        vst<V>(&ays[buf][w][lane * V], cy);
        lds_barrier();

        auto body = [&](auto E) {
            constexpr int e = decltype(E)::value;  // (j + p) & 1
            // red cells of plane p
            float ayd[V];  // ay of row j-1
            if (w > 0) vld<V>(&ays[buf][w - 1][lane * V], ayd);
            else {
#pragma unroll
                for (int c = 0; c < V; ++c) ayd[c] = exy[c];
            }
            const float axe = __shfl(cx[V - 1], lprev, 64);  // ax of cell i0-1
            float nr[H];
#pragma unroll
            for (int h = 0; h < H; ++h) nr[h] = 0.f;
            if (row_ok && p >= 0 && p < nk) {
                float xdn[H], xup[H], edge = 0.f;
                if constexpr (MODE != 1) {
                    if (w > 0) vld<H>(&xs[buf][w - 1][lane * H], xdn);
                    else {
#pragma unroll
                        for (int h = 0; h < H; ++h) xdn[h] = Rc<e>(exx, h);
                    }
                    if (w < kZRows - 1) vld<H>(&xs[buf][w + 1][lane * H], xup);
                    else {
#pragma unroll
                        for (int h = 0; h < H; ++h) xup[h] = Rc<e>(exx, h);
                    }
                    edge = e ? __shfl(xb0[H - 1], lprev, 64) : __shfl(xb0[0], lnext, 64);
                }
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const int i = i0 + 2 * h + 1 - e;
                    float s = 0.f;
                    if constexpr (MODE != 1) {
                        const bool zm = p > 0, xm = i > 0, xp = i + 1 < ni, zp = p + 1 < nk;
                        const float axl = e ? (h > 0 ? cx[2 * h - 1] : axe) : cx[2 * h];
                        const float a0 = zm ? sc * Rc<e>(zprev, h) : 0.f;
                        const float a1 = ym ? sc * Rc<e>(ayd, h) : 0.f;
                        const float a2 = xm ? sc * axl : 0.f;
                        const float a3 = xp ? sc * Rc<e>(cx, h) : 0.f;
                        const float a4 = yp ? sc * Rc<e>(cy, h) : 0.f;
                        const float a5 = zp ? sc * Rc<e>(cz, h) : 0.f;
                        const float xl = e ? (h > 0 ? xb0[h - 1] : edge) : xb0[h];
                        const float xr = e ? xb0[h] : (h + 1 < H ? xb0[h + 1] : edge);
                        const float x0 = zm ? xbm[h] : 0.f;
                        const float x1 = ym ? xdn[h] : 0.f;
                        const float x2 = xm ? xl : 0.f;
                        const float x3 = xp ? xr : 0.f;
                        const float x4 = yp ? xup[h] : 0.f;
                        const float x5 = zp ? xbp[h] : 0.f;
                        s = __builtin_fmaf(a0, x0, s);
                        s = __builtin_fmaf(a1, x1, s);
                        s = __builtin_fmaf(a2, x2, s);
                        s = __builtin_fmaf(a3, x3, s);
                        s = __builtin_fmaf(a4, x4, s);
                        s = __builtin_fmaf(a5, x5, s);
                    }
                    nr[h] = gs_value(Rc<e>(cb, h), s, sc * Rc<e>(cd, h));
                }
            }
            vst<H>(&rs[buf][w][lane * H], nr);

            // black cells of plane q = p-1 (same positions as this plane's red cells)
            const int q = p - 1;
            if (out_row && q >= k0 && q < k1) {
                float rdn[H], rup[H];
                vld<H>(&rs[buf ^ 1][w - 1][lane * H], rdn);
                vld<H>(&rs[buf ^ 1][w + 1][lane * H], rup);
                const float edge = e ? __shfl(nq[H - 1], lprev, 64) : __shfl(nq[0], lnext, 64);
                float o[V];
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const int i = i0 + 2 * h + 1 - e;
                    const float rl = e ? (h > 0 ? nq[h - 1] : edge) : nq[h];
                    const float rr = e ? nq[h] : (h + 1 < H ? nq[h + 1] : edge);
                    float sum = 0.f;
                    if (q > 0) sum = __builtin_fmaf(qzd[h], nqm[h], sum);
                    if (ym) sum = __builtin_fmaf(qyd[h], rdn[h], sum);
                    if (i > 0) sum = __builtin_fmaf(qxl[h], rl, sum);
                    if (i + 1 < ni) sum = __builtin_fmaf(qxr[h], rr, sum);
                    if (yp) sum = __builtin_fmaf(qy[h], rup[h], sum);
                    if (q + 1 < nk) sum = __builtin_fmaf(qz[h], nr[h], sum);
                    const float v = gs_value(qb[h], sum, qd[h]);
                    o[2 * h + 1 - e] = v;
                    o[2 * h + e] = nq[h];
                    if (dot) dacc += (Acc)v * (Acc)qb[h] + (Acc)nq[h] * (Acc)qbr[h];
                }
                vst<V>(xout + (q * nj + j) * ni + i0, o);
            }

            // black cells of plane p for the next step (c = 2h + e), then rotate
#pragma unroll
            for (int h = 0; h < H; ++h) {
                qb[h] = Bc<e>(cb, h);
                qbr[h] = Rc<e>(cb, h);
                qd[h] = sc * Bc<e>(cd, h);
                qxl[h] = sc * (e ? cx[2 * h] : (h > 0 ? cx[2 * h - 1] : axe));
                qxr[h] = sc * Bc<e>(cx, h);
                qy[h] = sc * Bc<e>(cy, h);
                qyd[h] = sc * Bc<e>(ayd, h);
                qz[h] = sc * Bc<e>(cz, h);
                qzd[h] = sc * Bc<e>(zprev, h);
                nqm[h] = nq[h];
                nq[h] = nr[h];
            }
            if constexpr (MODE != 1) {
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    xbm[h] = xb0[h];
                    xb0[h] = xbp[h];
                    xbp[h] = Bc<e>(nxx, h);  // plane p+2 has the parity of p
                }
            }
        };
        if ((j + p) & 1) body(I1()); else body(I0());
#pragma unroll
        for (int c = 0; c < V; ++c) {
            zprev[c] = cz[c];
            cb[c] = nb_[c]; cd[c] = nd[c]; cx[c] = nx[c]; cy[c] = ny[c]; cz[c] = nz[c];
            exy[c] = ney[c];
            if constexpr (MODE != 1) exx[c] = nex[c];
        }
    }
    if (dot) {
        const Acc t = wave_sum(dacc);
        if (lane == 0) part[w] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            Acc r = 0;
            for (int k = 0; k < kZRows; ++k) r += part[k];
            dot[blockIdx.x] = (double)r;
        }
    }
}

// ----------------------------------------------------------------------------
// b_c = R (b - A x), R = 1/8 per existing child (S/AlgebraicMultigrid.h:105-117,
// S/GeometricLevelGen.cpp:50-63). A 64 x 2 x 2 block computes one fine residual
// per thread (coalesced rows) and reduces each 2x2x2 aggregate: the x pair with
// a lane shuffle, the y/z pairs through LDS.
// ----------------------------------------------------------------------------
template <typename CT>
__device__ inline void residual_restrict_body(const Op<CT>& op, const float* __restrict__ b,
                                              const float* __restrict__ x, float* __restrict__ bc, int ci, int cj,
                                              int ck, Blk B) {
    __shared__ float part[4][32];
    const int i = B.x * 64 + threadIdx.x;
    const int j = B.y * 2 + threadIdx.y;
    const int k = B.z * 2 + threadIdx.z;
    float r = 0.f;
    if (i < op.ni && j < op.nj && k < op.nk) {
        const int c = (k * op.nj + j) * op.ni + i;
        r = 0.125f * (b[c] - __builtin_fmaf(op.D(c), x[c], offdiag_sum(op, x, c, i, j, k)));
    }
    r += __shfl_xor(r, 1, 64);
    const int row = threadIdx.y + 2 * threadIdx.z;
    if ((threadIdx.x & 1) == 0) part[row][threadIdx.x >> 1] = r;
    __syncthreads();
    if (row == 0 && (threadIdx.x & 1) == 0) {
        const int m = threadIdx.x >> 1;
        const int I = B.x * 32 + m, J = B.y, K = B.z;
        if (I < ci && J < cj && K < ck)
            bc[(K * cj + J) * ci + I] = (part[0][m] + part[1][m]) + (part[2][m] + part[3][m]);
    }
}
template <typename CT>
__global__ __launch_bounds__(256) void k_residual_restrict(Op<CT> op, const float* __restrict__ b,
                                                           const float* __restrict__ x, float* __restrict__ bc,
                                                           int ci, int cj, int ck) {
    residual_restrict_body(op, b, x, bc, ci, cj, ck, this_blk());
}

// ----------------------------------------------------------------------------
// Coarsest level (<= 4096 unknowns, S/GeometricLevelGen.h:46): all RB-GS sweeps
// in one workgroup, x in LDS, operator in registers. Each thread owns up to 2
// x-pairs (2m, 2m+1) of a row; a pair holds one cell of each colour, so every
// lane has work in both half sweeps. LDS holds x split by colour, slot
// colour * npairs + pair index, so the lanes of a wave touch consecutive words
// (no bank conflicts) where cell order would give them a stride of two.
// ----------------------------------------------------------------------------
constexpr int kCoarseMax = 4096;
constexpr int kCoarseThreads = 1024;
constexpr int kCoarsePairs = 2;  // pairs per thread: up to 2048 pairs

// LDS slot of cell (i, j, k): its colour's half, then row-major pair order
__device__ inline int coarse_slot(int i, int j, int k, int nj, int ppr, int npairs) {
    return ((i + j + k) & 1) * npairs + (k * nj + j) * ppr + (i >> 1);
}

template <typename CT>
__device__ inline void coarse_cell(const Op<CT>& op, const float* __restrict__ b, int unit, int i, int j, int k,
                                   int ppr, int npairs, int& cell, int& slot, float& d, float& bb, float (&a)[6],
                                   int (&nb)[6]) {
    const int ni = op.ni, nj = op.nj, nk = op.nk, sl = ni * nj;
    if (i >= ni) return;
    const int c = (k * nj + j) * ni + i;
    cell = c;
    slot = coarse_slot(i, j, k, nj, ppr, npairs);
    d = op.D(c);
    bb = unit >= 0 ? (c == unit ? 1.f : 0.f) : b[c];
    if (k > 0) { nb[0] = coarse_slot(i, j, k - 1, nj, ppr, npairs); a[0] = op.AZ(c - sl); }
    if (j > 0) { nb[1] = coarse_slot(i, j - 1, k, nj, ppr, npairs); a[1] = op.AY(c - ni); }
    if (i > 0) { nb[2] = coarse_slot(i - 1, j, k, nj, ppr, npairs); a[2] = op.AX(c - 1); }
    if (i + 1 < ni) { nb[3] = coarse_slot(i + 1, j, k, nj, ppr, npairs); a[3] = op.AX(c); }
    if (j + 1 < nj) { nb[4] = coarse_slot(i, j + 1, k, nj, ppr, npairs); a[4] = op.AY(c); }
    if (k + 1 < nk) { nb[5] = coarse_slot(i, j, k + 1, nj, ppr, npairs); a[5] = op.AZ(c); }
}

// Each cell keeps 1/diag (0 where the diagonal is 0, as gs_value) in a register:
// the 200 sweeps then multiply instead of dividing, which halves the VALU work of
// an update; the result differs from the divided form by rounding only.
// tab != nullptr: tabulation mode. Workgroup u solves with b = e_u and writes
// the result to row u of tab (tab[u*n + r] = (RBGS^sweeps e_u)[r]); the solve
// from a zero guess is linear in b, so y = tab^T b reproduces it (k_coarse_apply).
// XREG (the host sets it when a row's pairs, (ni + 1) / 2 of them, divide 16):
// a row's pairs then sit on consecutive lanes of one 16-lane DPP row, so both
// x-neighbours of a cell come from registers: one from the thread's own other
// cell, one from the next or previous lane through a DPP row shift. Only the y
// and z neighbours are read from LDS; the terms and their order are unchanged.
__device__ inline float dpp_from_prev(float v) {  // lane l - 1 of the 16-lane row (0 at its start)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, false));
}
__device__ inline float dpp_from_next(float v) {  // lane l + 1 of the 16-lane row (0 at its end)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xf, 0xf, false));
}

template <typename CT, bool XREG = false>
__global__ __launch_bounds__(kCoarseThreads) void k_coarsest(Op<CT> op, const float* __restrict__ b,
                                                             float* __restrict__ x, int sweeps,
                                                             float* __restrict__ tab) {
    __shared__ float xs[kCoarseMax + 1];
    const int unit = tab ? (int)blockIdx.x : -1;
    const int ni = op.ni, nj = op.nj, nk = op.nk;
    const int ppr = (ni + 1) / 2;
    const int npairs = ppr * nj * nk;
    // [q][colour]: colour 1 = odd (i+j+k), updated first; cd holds 1/diag after set-up
    float cd[kCoarsePairs][2], cb[kCoarsePairs][2], ca[kCoarsePairs][2][6];
    int nbr[kCoarsePairs][2][6], cell[kCoarsePairs][2], slot[kCoarsePairs][2];
    float xv[kCoarsePairs][2];  // current x of the thread's cells (XREG)
    int ev[kCoarsePairs];       // which of the pair's cells has the even i (XREG)
#pragma unroll
    for (int q = 0; q < kCoarsePairs; ++q) {
#pragma unroll
        for (int col = 0; col < 2; ++col) {
            cell[q][col] = -1;
            slot[q][col] = kCoarseMax;
            xv[q][col] = 0.f;
            ev[q] = 0;
            cd[q][col] = 0.f;
            cb[q][col] = 0.f;
#pragma unroll
            for (int e = 0; e < 6; ++e) {
                nbr[q][col][e] = kCoarseMax;  // zero pad
                ca[q][col][e] = 0.f;
            }
        }
        const int P = threadIdx.x + q * kCoarseThreads;
        if (P < npairs) {
            const int row = P / ppr, m = P - row * ppr;
            const int j = row % nj, k = row / nj;
            // cell 2m+e is odd when e == 1 - ((j+k)&1)
            const int eodd = ((j + k) & 1) ? 0 : 1;
            ev[q] = eodd ? 0 : 1;
            coarse_cell(op, b, unit, 2 * m + eodd, j, k, ppr, npairs, cell[q][1], slot[q][1], cd[q][1], cb[q][1],
                        ca[q][1], nbr[q][1]);
            coarse_cell(op, b, unit, 2 * m + 1 - eodd, j, k, ppr, npairs, cell[q][0], slot[q][0], cd[q][0],
                        cb[q][0], ca[q][0], nbr[q][0]);
        }
#pragma unroll
        for (int col = 0; col < 2; ++col) cd[q][col] = cd[q][col] != 0.f ? 1.f / cd[q][col] : 0.f;
#pragma unroll
        for (int col = 0; col < 2; ++col)
            if (cell[q][col] >= 0) xs[slot[q][col]] = 0.f;
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
                if constexpr (XREG) {
                    const float xe = ev[q] ? xv[q][1] : xv[q][0], xo = ev[q] ? xv[q][0] : xv[q][1];
                    const float xo_prev = dpp_from_prev(xo), xe_next = dpp_from_next(xe);
                    const bool even = col == ev[q];
                    s += ca[q][col][0] * xs[nbr[q][col][0]];
                    s += ca[q][col][1] * xs[nbr[q][col][1]];
                    s += ca[q][col][2] * (even ? xo_prev : xe);  // -x
                    s += ca[q][col][3] * (even ? xo : xe_next);  // +x
                    s += ca[q][col][4] * xs[nbr[q][col][4]];
                    s += ca[q][col][5] * xs[nbr[q][col][5]];
                } else {
#pragma unroll
                    for (int e = 0; e < 6; ++e) s += ca[q][col][e] * xs[nbr[q][col][e]];
                }
                nv[q] = (cb[q][col] - s) * cd[q][col];
            }
#pragma unroll
            for (int q = 0; q < kCoarsePairs; ++q)
                if (cell[q][col] >= 0) {
                    xs[slot[q][col]] = nv[q];
                    if constexpr (XREG) xv[q][col] = nv[q];
                }
            __syncthreads();
        }
    }
#pragma unroll
    for (int q = 0; q < kCoarsePairs; ++q)
#pragma unroll
        for (int col = 0; col < 2; ++col)
            if (cell[q][col] >= 0) {
                if (tab) tab[(size_t)unit * (op.ni * op.nj * op.nk) + cell[q][col]] = xs[slot[q][col]];
                else x[cell[q][col]] = xs[slot[q][col]];
            }
}

// Coarsest level whose pairs exceed the register-resident kernel's capacity
// (npairs > kCoarsePairs * kCoarseThreads, e.g. 15 x 17 x 16 with 8 pairs per row):
// the same RB-GS sweeps (odd colour first, reference CSR term order, division by
// the diagonal as gs_value) with x in LDS at its cell index and the operator read
// from memory; each thread strides over the level's cells. tab as k_coarsest.
template <typename CT>
__global__ __launch_bounds__(kCoarseThreads) void k_coarsest_general(Op<CT> op, const float* __restrict__ b,
                                                                     float* __restrict__ x, int sweeps,
                                                                     float* __restrict__ tab) {
    __shared__ float xs[kCoarseMax];
    const int unit = tab ? (int)blockIdx.x : -1;
    const int ni = op.ni, nj = op.nj, n = op.ni * op.nj * op.nk;
    for (int c = threadIdx.x; c < n; c += kCoarseThreads) xs[c] = 0.f;
    __syncthreads();
    for (int it = 0; it < sweeps; ++it) {
        for (int col = 1; col >= 0; --col) {
            for (int c = threadIdx.x; c < n; c += kCoarseThreads) {
                int i, j, k;
                unflat32((unsigned)c, (unsigned)ni, (unsigned)nj, i, j, k);
                if (((i + j + k) & 1) != col) continue;
                const float bc = unit >= 0 ? (c == unit ? 1.f : 0.f) : b[c];
                xs[c] = gs_value(bc, offdiag_sum(op, xs, c, i, j, k), op.D(c));
            }
            __syncthreads();
        }
    }
    for (int c = threadIdx.x; c < n; c += kCoarseThreads) {
        if (tab) tab[(size_t)unit * n + c] = xs[c];
        else x[c] = xs[c];
    }
}

// Whether the register-resident coarsest kernel covers a level's pairs.
__host__ __device__ inline bool coarse_fits_registers(int ni, int nj, int nk) {
    return ((ni + 1) / 2) * nj * nk <= kCoarsePairs * kCoarseThreads;
}

// y = tab^T b in two passes (64-column chunks, then a fixed-order sum of the
// chunk partials): deterministic and spread over the whole GPU.
constexpr int kCoarseChunk = 64;
__device__ inline void coarse_apply_partial_body(const float* __restrict__ tab, const float* __restrict__ b,
                                                 float* __restrict__ part, int n, Blk B) {
    __shared__ float bs[kCoarseChunk];
    const int r = B.x * kBlock + threadIdx.x;
    const int c0 = B.y * kCoarseChunk;
    const int cn = min(kCoarseChunk, n - c0);
    if (threadIdx.x < cn) bs[threadIdx.x] = b[c0 + threadIdx.x];
    __syncthreads();
    if (r >= n) return;
    float acc = 0.f;
#pragma unroll 16
    for (int q = 0; q < cn; ++q) acc += tab[(size_t)(c0 + q) * n + r] * bs[q];
    part[(size_t)B.y * n + r] = acc;
}
static __global__ __launch_bounds__(kBlock) void k_coarse_apply_partial(const float* __restrict__ tab,
                                                                 const float* __restrict__ b,
                                                                 float* __restrict__ part, int n) {
    coarse_apply_partial_body(tab, b, part, n, this_blk());
}
__device__ inline void coarse_apply_reduce_body(const float* __restrict__ part, float* __restrict__ y, int n,
                                                int nchunks, Blk B) {
    const int r = B.x * kBlock + threadIdx.x;
    if (r >= n) return;
    // the loads of 16 chunks are issued together; the sum keeps the chunk order
    float s = 0.f;
#pragma unroll 16
    for (int q = 0; q < nchunks; ++q) s += part[(size_t)q * n + r];
    y[r] = s;
}
static __global__ __launch_bounds__(kBlock) void k_coarse_apply_reduce(const float* __restrict__ part, float* __restrict__ y,
                                                                int n, int nchunks) {
    coarse_apply_reduce_body(part, y, n, nchunks, this_blk());
}

// ----------------------------------------------------------------------------
// Closed-form Galerkin coarsening A_c = 0.5 R A P (R = 1/8, P = injection):
//   coupling(I, I+e) = 1/16 * sum of the fine couplings across the aggregate face
//   leak(I)          = 1/16 * sum of the children's leaks
//   diag(I)          = leak(I) + sum of the 6 coarse couplings' magnitudes
// (equal to 1/16 * sum_{i,k in I} A(i,k); written this way no term cancels).
// ----------------------------------------------------------------------------
static __global__ __launch_bounds__(kBlock) void k_coarsen(int ni, int nj, int nk, const float* __restrict__ ax,
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
static __global__ __launch_bounds__(kBlock) void k_leak_from_diag(int ni, int nj, int nk, const float* __restrict__ d,
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
static __global__ __launch_bounds__(kBlock) void k_to_half(unsigned n, const float* __restrict__ src, __half* __restrict__ dst,
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
    double energy;  // E(x0) = x0'A x0 / 2 - b'x0 of a warm start
};

// q = A s and sum(s.q); 3D launch (64 x 4 threads, kMatvecKZ planes per block).
template <typename CT, typename Acc>
__device__ inline void matvec_dot_body(const Op<CT>& op, const float* __restrict__ s, float* __restrict__ q,
                                       double* __restrict__ partial, Blk B) {
    const int i = B.x * 64 + threadIdx.x, j = B.y * 4 + threadIdx.y;
    const int kb = B.z * kMatvecKZ;
    const int ni = op.ni, nj = op.nj, nk = op.nk, sl = ni * nj;
    Acc acc = 0;
    if (i < ni && j < nj) {
        for (int k = kb; k < min(kb + kMatvecKZ, nk); ++k) {
            const int c = (k * nj + j) * ni + i;
            // reference multiply order: -z, -y, -x, diag, +x, +y, +z
            float y = 0.f;
            if (k > 0) y = __builtin_fmaf(op.AZ(c - sl), s[c - sl], y);
            if (j > 0) y = __builtin_fmaf(op.AY(c - ni), s[c - ni], y);
            if (i > 0) y = __builtin_fmaf(op.AX(c - 1), s[c - 1], y);
            y = __builtin_fmaf(op.D(c), s[c], y);
            if (i + 1 < ni) y = __builtin_fmaf(op.AX(c), s[c + 1], y);
            if (j + 1 < nj) y = __builtin_fmaf(op.AY(c), s[c + ni], y);
            if (k + 1 < nk) y = __builtin_fmaf(op.AZ(c), s[c + sl], y);
            q[c] = y;
            acc += (Acc)s[c] * (Acc)y;
        }
    }
    Acc t = block_sum(acc);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[(B.z * B.ny + B.y) * B.nx + B.x] = (double)t;
}
template <typename CT, typename Acc>
__global__ __launch_bounds__(kBlock) void k_matvec_dot(Op<CT> op, const float* __restrict__ s, float* __restrict__ q,
                                                       double* __restrict__ partial) {
    matvec_dot_body<CT, Acc>(op, s, q, partial, this_blk());
}

// Row-wave 7-point stencil for levels whose rows are 64*V cells (V = 2 or 4),
// organised like k_rbgs_zrow: a wave owns one x row, a 512-thread workgroup
// kSRows consecutive rows (two workgroups per compute unit: one marches while the
// other waits at its barrier), marching up a z chunk with the z-neighbours in
// registers, the y-neighbours through LDS (double-buffered, one barrier per
// plane; the first and last wave also load the row just outside the tile) and
// the x-neighbours by lane shuffles.
//   KIND 0: out = A s and a per-workgroup sum(s.out)      (k_matvec_dot)
//   KIND 1: out = R (b - A s) on the next coarser level   (k_residual_restrict)
// Same terms, same order and the same fused multiply-adds as those kernels. The
// restriction sums each 2x2x2 aggregate as they do: the x pair inside the lane,
// the y pair (rows 2J, 2J+1 = waves w, w+1) through LDS one plane later, the z
// pair in registers; z chunks start on even planes.
constexpr int kSRows = 8;  // rows per workgroup of k_stencil_row (even: restriction row pairs)
template <int V, int KIND, typename Acc, typename CT = float>
__global__ __launch_bounds__(64 * kSRows) void k_stencil_row(Op<CT> op, const float* __restrict__ b,
                                                      const float* __restrict__ s, float* __restrict__ out, int ci,
                                                      int cj, int tiles_y, int zchunk, double* __restrict__ partial) {
    constexpr int H = V / 2;
    __shared__ __align__(16) float ss[2][kSRows][64 * V];   // s of plane k, per row
    __shared__ __align__(16) float ays[2][kSRows][64 * V];  // ay of plane k, per row
    __shared__ __align__(16) float ps[2][kSRows][64 * H];   // restriction: x-pair sums of plane k
    __shared__ Acc part[kSRows];
    const int ni = op.ni, nj = op.nj, nk = op.nk;
    const float sc = op.scale;
    const int nb = gridDim.x, L = blockIdx.x, xcd = L % 8;
    const int lg = xcd * (nb / 8) + min(xcd, nb % 8) + L / 8;
    const int ty = lg % tiles_y, zch = lg / tiles_y;
    const int j0 = ty * kSRows, k0 = zch * zchunk, k1 = min(k0 + zchunk, nk);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int j = j0 + w;
    const bool row_ok = j < nj;
    const bool ym = j > 0, yp = j + 1 < nj;
    const int jc = min(j, nj - 1);
    const bool ex_lo = w == 0, ex_hi = w == kSRows - 1;
    const int jx = ex_lo ? max(j - 1, 0) : min(j + 1, nj - 1);
    const int i0 = lane * V;
    const int lprev = (lane + 63) & 63, lnext = (lane + 1) & 63;
    Acc dacc = 0;
    auto at = [&](int jj, int k) { return (min(max(k, 0), nk - 1) * nj + jj) * ni + i0; };

    float sm[V], s0[V], sp[V];                 // s of row j at planes k-1, k, k+1
    float cd[V], cx[V], cy[V], cz[V], cb[V];   // fields of plane k (cb: KIND 1)
    float zprev[V];                            // az of plane k-1
    float exs[V], exy[V];                      // row outside the tile: s (and ay below) at plane k
    float xpl[H], ylo[H];                      // KIND 1: own x-pair sums of plane k-1; y-pair sums of the even plane
    {
        vld<V>(s + at(jc, k0 - 1), sm);
        vld<V>(s + at(jc, k0), s0);
        vld<V>(s + at(jc, k0 + 1), sp);
        const int o = at(jc, k0);
        vld<V>(op.d + o, cd); vld<V>(op.ax + o, cx); vld<V>(op.ay + o, cy); vld<V>(op.az + o, cz);
        if constexpr (KIND == 1) vld<V>(b + o, cb);
        vld<V>(op.az + at(jc, k0 - 1), zprev);
        if (ex_lo || ex_hi) vld<V>(s + at(jx, k0), exs);
        if (ex_lo) vld<V>(op.ay + at(jx, k0), exy);
#pragma unroll
        for (int h = 0; h < H; ++h) xpl[h] = ylo[h] = 0.f;
    }
    for (int k = k0; k <= k1; ++k) {
        const int buf = k & 1;
        float nd[V], nx[V], ny[V], nz[V], nbv[V], nsp[V], nexs[V], nexy[V];
        if (k < k1) {
            const int o = at(jc, k + 1);
            vld<V>(op.d + o, nd); vld<V>(op.ax + o, nx); vld<V>(op.ay + o, ny); vld<V>(op.az + o, nz);
            if constexpr (KIND == 1) vld<V>(b + o, nbv);
            vld<V>(s + at(jc, k + 2), nsp);
            if (ex_lo || ex_hi) vld<V>(s + at(jx, k + 1), nexs);
            if (ex_lo) vld<V>(op.ay + at(jx, k + 1), nexy);
            vst<V>(&ss[buf][w][lane * V], s0);
            vst<V>(&ays[buf][w][lane * V], cy);
        }
        lds_barrier();
        if constexpr (KIND == 1) {
            // plane k-1: the even wave of each row pair adds its partner's x-pair sums
            if (k > k0 && (w & 1) == 0 && row_ok) {
                const int kq = k - 1;
                float other[H], out_v[H];
                vld<H>(&ps[buf ^ 1][w + 1][lane * H], other);
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const float ys = xpl[h] + other[h];
                    if (kq & 1) out_v[h] = ylo[h] + ys;
                    else if (kq == nk - 1) out_v[h] = ys + (0.f + 0.f);
                    else ylo[h] = ys;
                }
                if ((kq & 1) || kq == nk - 1) vst<H>(out + ((kq >> 1) * cj + (j >> 1)) * ci + lane * H, out_v);
            }
        }
        if (k < k1) {
            float xp[H];
#pragma unroll
            for (int h = 0; h < H; ++h) xp[h] = 0.f;
            if (row_ok) {
                float sdn[V], sup[V], ayd[V];
                if (w > 0) {
                    vld<V>(&ss[buf][w - 1][lane * V], sdn);
                    vld<V>(&ays[buf][w - 1][lane * V], ayd);
                } else {
#pragma unroll
                    for (int c = 0; c < V; ++c) { sdn[c] = exs[c]; ayd[c] = exy[c]; }
                }
                if (w < kSRows - 1) vld<V>(&ss[buf][w + 1][lane * V], sup);
                else {
#pragma unroll
                    for (int c = 0; c < V; ++c) sup[c] = exs[c];
                }
                const float sl_e = __shfl(s0[V - 1], lprev, 64);
                const float sr_e = __shfl(s0[0], lnext, 64);
                const float ax_e = __shfl(cx[V - 1], lprev, 64);
                float y[V];
#pragma unroll
                for (int c = 0; c < V; ++c) {
                    const int i = i0 + c;
                    float t = 0.f;
                    if (k > 0) t = __builtin_fmaf(sc * zprev[c], sm[c], t);
                    if (ym) t = __builtin_fmaf(sc * ayd[c], sdn[c], t);
                    if (i > 0) t = __builtin_fmaf(sc * (c > 0 ? cx[c - 1] : ax_e), c > 0 ? s0[c - 1] : sl_e, t);
                    if constexpr (KIND == 0) t = __builtin_fmaf(sc * cd[c], s0[c], t);
                    if (i + 1 < ni) t = __builtin_fmaf(sc * cx[c], c + 1 < V ? s0[c + 1] : sr_e, t);
                    if (yp) t = __builtin_fmaf(sc * cy[c], sup[c], t);
                    if (k + 1 < nk) t = __builtin_fmaf(sc * cz[c], sp[c], t);
                    if constexpr (KIND == 0) {
                        y[c] = t;
                        if (partial) dacc += (Acc)s0[c] * (Acc)t;
                    } else {
                        y[c] = 0.125f * (cb[c] - __builtin_fmaf(sc * cd[c], s0[c], t));
                    }
                }
                if constexpr (KIND == 0) vst<V>(out + (k * nj + j) * ni + i0, y);
                else {
#pragma unroll
                    for (int h = 0; h < H; ++h) xp[h] = y[2 * h] + y[2 * h + 1];
                }
            }
            if constexpr (KIND == 1) {
                vst<H>(&ps[buf][w][lane * H], xp);
#pragma unroll
                for (int h = 0; h < H; ++h) xpl[h] = xp[h];
            }
#pragma unroll
            for (int c = 0; c < V; ++c) {
                zprev[c] = cz[c];
                sm[c] = s0[c]; s0[c] = sp[c]; sp[c] = nsp[c];
                cd[c] = nd[c]; cx[c] = nx[c]; cy[c] = ny[c]; cz[c] = nz[c];
                if constexpr (KIND == 1) cb[c] = nbv[c];
                exs[c] = nexs[c]; exy[c] = nexy[c];
            }
        }
    }
    if constexpr (KIND == 0) {
        if (partial) {
            const Acc t = wave_sum(dacc);
            if (lane == 0) part[w] = t;
            __syncthreads();
            if (threadIdx.x == 0) {
                Acc r = 0;
                for (int q = 0; q < kSRows; ++q) r += part[q];
                partial[blockIdx.x] = (double)r;
            }
        }
    }
}

// FP64 true residual r = b - A x of the mixed-precision refinement. The level-0
// couplings are exact in FP32 (each is one FP32 face term); the diagonal and
// the right-hand side come in FP64 when the assembly provided them (d64, b64),
// else from the FP32 arrays. x is the FP64 iterate. Writes r rounded to FP32
// and partial max|r|.
// r = b - A x in FP64 (FP64 iterate, FP64 diagonal/rhs when given), partial
// max|r|. A thread marches kTrueResKZ planes up z with x and the z coupling of
// the plane below kept in registers; the terms are summed in the order
// d x, -z, -y, -x, +x, +y, +z. epart (optional): per-block sums of x (b + r)
// for the energy of x: E(x) = x'Ax/2 - b'x = -(x.b + x.r)/2 with r = b - A x.
constexpr int kTrueResKZ = 16;
// Block of 64 x kTrueResTY cells: the rows just outside it, read for the y
// neighbours, are a fraction 2 / kTrueResTY of the x traffic.
constexpr int kTrueResTY = 8;
inline dim3 true_res_grid(int ni, int nj, int nk) {
    return dim3((ni + 63) / 64, (nj + kTrueResTY - 1) / kTrueResTY, (nk + kTrueResKZ - 1) / kTrueResKZ);
}
// With gram given it also accumulates the pressure history's new Gram row for x:
// partial sums of x.Ax, p1.Ax and p2.Ax (p1 / p2 may be null) at gram[q * nblocks + block].
static __global__ __launch_bounds__(64 * kTrueResTY) void k_true_residual64(Op<float> op, const double* __restrict__ d64,
                                                            const double* __restrict__ b64,
                                                            const float* __restrict__ b32,
                                                            const double* __restrict__ x, float* __restrict__ r,
                                                            double* __restrict__ partial,
                                                            double* __restrict__ gram = nullptr,
                                                            const double* __restrict__ p1 = nullptr,
                                                            const double* __restrict__ p2 = nullptr,
                                                            double* __restrict__ epart = nullptr) {
    double g0 = 0, g1 = 0, g2 = 0, e = 0;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * kTrueResTY + threadIdx.y;
    const int kb = blockIdx.z * kTrueResKZ;
    const int ni = op.ni, nj = op.nj, nk = op.nk, sl = ni * nj;
    double m = 0;
    if (i < ni && j < nj && kb < nk) {
        const int ke = min(kb + kTrueResKZ, nk);
        int c = (kb * nj + j) * ni + i;
        double xm = kb > 0 ? x[c - sl] : 0.0, x0 = x[c];
        float azm = kb > 0 ? op.AZ(c - sl) : 0.f;
        // the fields of plane k are loaded while plane k - 1 computes (zero where a
        // neighbour is outside the grid, whose term is skipped)
        struct Plane {
            double d, b, xym, xxm, xxp, xyp;
            float aym, axm, axp, ayp;
        };
        auto load = [&](int cc, Plane& q) {
            q.d = d64 ? d64[cc] : (double)op.D(cc);
            q.b = b64 ? b64[cc] : (double)b32[cc];
            q.aym = j > 0 ? op.AY(cc - ni) : 0.f;
            q.xym = j > 0 ? x[cc - ni] : 0.0;
            q.axm = i > 0 ? op.AX(cc - 1) : 0.f;
            q.xxm = i > 0 ? x[cc - 1] : 0.0;
            q.axp = i + 1 < ni ? op.AX(cc) : 0.f;
            q.xxp = i + 1 < ni ? x[cc + 1] : 0.0;
            q.ayp = j + 1 < nj ? op.AY(cc) : 0.f;
            q.xyp = j + 1 < nj ? x[cc + ni] : 0.0;
        };
        Plane cur, nxt;
        load(c, cur);
        for (int k = kb; k < ke; ++k, c += sl) {
            const double xp = k + 1 < nk ? x[c + sl] : 0.0;
            const float az0 = k + 1 < nk ? op.AZ(c) : 0.f;
            if (k + 1 < ke) load(c + sl, nxt);
            double y = cur.d * x0;
            if (k > 0) y += (double)azm * xm;
            if (j > 0) y += (double)cur.aym * cur.xym;
            if (i > 0) y += (double)cur.axm * cur.xxm;
            if (i + 1 < ni) y += (double)cur.axp * cur.xxp;
            if (j + 1 < nj) y += (double)cur.ayp * cur.xyp;
            if (k + 1 < nk) y += (double)az0 * xp;
            const double bv = cur.b;
            const double rv = bv - y;
            if (r) r[c] = (float)rv;
            // energy of x: x (b + r) with r as stored in FP32
            if (epart) e += x0 * (bv + (double)(float)rv);
            if (gram) {
                g0 += x0 * y;
                if (p1) g1 += p1[c] * y;
                if (p2) g2 += p2[c] * y;
            }
            m = fmax(m, fabs(rv));
            xm = x0;
            x0 = xp;
            azm = az0;
            cur = nxt;
        }
    }
    double t = block_max<double, 64 * kTrueResTY>(m);
    const unsigned blk = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blk] = t;
    if (epart) {
        const double se = block_sum<double, 64 * kTrueResTY>(e);
        if (threadIdx.x == 0 && threadIdx.y == 0) epart[blk] = se;
    }
    if (gram) {
        const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
        const double s0 = block_sum<double, 64 * kTrueResTY>(g0), s1 = block_sum<double, 64 * kTrueResTY>(g1),
                     s2 = block_sum<double, 64 * kTrueResTY>(g2);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            gram[blk] = s0;
            gram[nb + blk] = s1;
            gram[2 * nb + blk] = s2;
        }
    }
}

// Row-wave form of k_true_residual64 for levels whose rows are 64*V cells (V = 2
// or 4), organised like k_stencil_row: a wave owns one x row (lane l: cells
// V*l .. V*l+V-1), a 512-thread workgroup kTrRows consecutive rows, marching up a
// z chunk with the z-neighbours in registers, the y-neighbours through LDS
// (double-buffered, one barrier per plane; the first and last wave also load the
// row just outside the tile) and the x-neighbours by lane shuffles; the next
// plane's fields are loaded while the current one computes. Each cell sums the
// same terms in the same order as k_true_residual64 (r, max|r| exactly as it);
// the per-workgroup energy and Gram partial sums cover other cell sets.
constexpr int kTrRows = 8;
template <int V, bool GRAM>  // GRAM: the Gram partials (gram, p1, p2) are accumulated
__global__ __launch_bounds__(64 * kTrRows) void k_true_residual_row(
    Op<float> op, const double* __restrict__ d64, const double* __restrict__ b64, const float* __restrict__ b32,
    const double* __restrict__ x, float* __restrict__ r, double* __restrict__ partial, int tiles_y, int zchunk,
    double* __restrict__ gram, const double* __restrict__ p1, const double* __restrict__ p2,
    double* __restrict__ epart) {
    __shared__ __align__(16) double xs[2][kTrRows][64 * V];  // x of plane k, per row
    __shared__ __align__(16) float ays[2][kTrRows][64 * V];  // ay of plane k, per row
    const int ni = op.ni, nj = op.nj, nk = op.nk;
    const int nb = gridDim.x, L = blockIdx.x, xcd = L % 8;
    const int lg = xcd * (nb / 8) + min(xcd, nb % 8) + L / 8;
    const int ty = lg % tiles_y, zch = lg / tiles_y;
    const int j0 = ty * kTrRows, k0 = zch * zchunk, k1 = min(k0 + zchunk, nk);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int j = j0 + w;
    const bool row_ok = j < nj;
    const bool ym = j > 0, yp = j + 1 < nj;
    const int jc = min(j, nj - 1);
    const bool ex_lo = w == 0, ex_hi = w == kTrRows - 1;
    const int jx = ex_lo ? max(j - 1, 0) : min(j + 1, nj - 1);
    const int i0 = lane * V;
    const int lprev = (lane + 63) & 63, lnext = (lane + 1) & 63;
    auto at = [&](int jj, int k) { return (min(max(k, 0), nk - 1) * nj + jj) * ni + i0; };
    auto vldd = [&](const double* p, double (&v)[V]) {
#pragma unroll
        for (int c = 0; c < V; c += 2) {
            const double2 t = *reinterpret_cast<const double2*>(p + c);
            v[c] = t.x;
            v[c + 1] = t.y;
        }
    };
    // fields of a plane: diagonal and right-hand side widened to FP64
    auto load_db = [&](int o, double (&d)[V], double (&bb)[V]) {
        if (d64) {
            vldd(d64 + o, d);
        } else {
            float t[V];
            vld<V>(op.d + o, t);
#pragma unroll
            for (int c = 0; c < V; ++c) d[c] = (double)t[c];
        }
        if (b64) {
            vldd(b64 + o, bb);
        } else {
            float t[V];
            vld<V>(b32 + o, t);
#pragma unroll
            for (int c = 0; c < V; ++c) bb[c] = (double)t[c];
        }
    };
    double xm[V], x0[V], xp[V];        // x of row j at planes k-1, k, k+1
    double cd[V], cb[V], q1[V], q2[V];  // plane k: diagonal, rhs, history directions
    float cx[V], cy[V], cz[V], zprev[V];
    double exs[V];                      // row outside the tile: x at plane k
    float exy[V];                       // and (first wave) its ay
    double m = 0, e = 0, g0 = 0, g1 = 0, g2 = 0;
    {
        vldd(x + at(jc, k0 - 1), xm);
        vldd(x + at(jc, k0), x0);
        vldd(x + at(jc, k0 + 1), xp);
        const int o = at(jc, k0);
        load_db(o, cd, cb);
        vld<V>(op.ax + o, cx);
        vld<V>(op.ay + o, cy);
        vld<V>(op.az + o, cz);
        vld<V>(op.az + at(jc, k0 - 1), zprev);
        if (GRAM && p1) vldd(p1 + o, q1);
        if (GRAM && p2) vldd(p2 + o, q2);
        if (ex_lo || ex_hi) vldd(x + at(jx, k0), exs);
        if (ex_lo) vld<V>(op.ay + at(jx, k0), exy);
    }
    for (int k = k0; k < k1; ++k) {
        const int buf = k & 1;
        double nd[V], nbv[V], nq1[V], nq2[V], nxp[V], nexs[V];
        float nx[V], ny[V], nz[V], nexy[V];
        if (k + 1 < k1) {
            const int o = at(jc, k + 1);
            load_db(o, nd, nbv);
            vld<V>(op.ax + o, nx);
            vld<V>(op.ay + o, ny);
            vld<V>(op.az + o, nz);
            if (GRAM && p1) vldd(p1 + o, nq1);
            if (GRAM && p2) vldd(p2 + o, nq2);
            vldd(x + at(jc, k + 2), nxp);
            if (ex_lo || ex_hi) vldd(x + at(jx, k + 1), nexs);
            if (ex_lo) vld<V>(op.ay + at(jx, k + 1), nexy);
        }
#pragma unroll
        for (int c = 0; c < V; ++c) {
            xs[buf][w][lane * V + c] = x0[c];
            ays[buf][w][lane * V + c] = cy[c];
        }
        lds_barrier();
        if (row_ok) {
            double xdn[V], xup[V];
            float ayd[V];
#pragma unroll
            for (int c = 0; c < V; ++c) {
                xdn[c] = w > 0 ? xs[buf][w - 1][lane * V + c] : exs[c];
                ayd[c] = w > 0 ? ays[buf][w - 1][lane * V + c] : exy[c];
                xup[c] = w < kTrRows - 1 ? xs[buf][w + 1][lane * V + c] : exs[c];
            }
            const double xl_e = __shfl(x0[V - 1], lprev, 64);
            const double xr_e = __shfl(x0[0], lnext, 64);
            const float axl_e = __shfl(cx[V - 1], lprev, 64);
#pragma unroll
            for (int c = 0; c < V; ++c) {
                const int i = i0 + c;
                double y = cd[c] * x0[c];
                if (k > 0) y += (double)zprev[c] * xm[c];
                if (ym) y += (double)ayd[c] * xdn[c];
                if (i > 0) y += (double)(c > 0 ? cx[c - 1] : axl_e) * (c > 0 ? x0[c - 1] : xl_e);
                if (i + 1 < ni) y += (double)cx[c] * (c + 1 < V ? x0[c + 1] : xr_e);
                if (yp) y += (double)cy[c] * xup[c];
                if (k + 1 < nk) y += (double)cz[c] * xp[c];
                const double bv = cb[c];
                const double rv = bv - y;
                if (r) r[(k * nj + j) * ni + i] = (float)rv;
                if (epart) e += x0[c] * (bv + (double)(float)rv);
                if (GRAM && gram) {
                    g0 += x0[c] * y;
                    if (p1) g1 += q1[c] * y;
                    if (p2) g2 += q2[c] * y;
                }
                m = fmax(m, fabs(rv));
            }
        }
        if (k + 1 < k1) {
#pragma unroll
            for (int c = 0; c < V; ++c) {
                zprev[c] = cz[c];
                xm[c] = x0[c];
                x0[c] = xp[c];
                xp[c] = nxp[c];
                cd[c] = nd[c];
                cb[c] = nbv[c];
                cx[c] = nx[c];
                cy[c] = ny[c];
                cz[c] = nz[c];
                if (GRAM && p1) q1[c] = nq1[c];
                if (GRAM && p2) q2[c] = nq2[c];
                exs[c] = nexs[c];
                exy[c] = nexy[c];
            }
        }
    }
    const double t = block_max<double, 64 * kTrRows>(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
    if (epart) {
        const double se = block_sum<double, 64 * kTrRows>(e);
        if (threadIdx.x == 0) epart[blockIdx.x] = se;
    }
    if (GRAM && gram) {
        const double s0 = block_sum<double, 64 * kTrRows>(g0), s1 = block_sum<double, 64 * kTrRows>(g1),
                     s2 = block_sum<double, 64 * kTrRows>(g2);
        if (threadIdx.x == 0) {
            gram[blockIdx.x] = s0;
            gram[nb + blockIdx.x] = s1;
            gram[2 * nb + blockIdx.x] = s2;
        }
    }
}

// Residual r = b - A x of an FP64 guess (the terms and FP64 arithmetic of
// k_true_residual64, rounded to FP32 per cell), restricted by 4 per axis for the
// network's input without storing r: crhs = mean of r over each 4^3 block, csolid =
// the fraction of its cells with diagonal 0, cliquid = 1 - csolid. Block 256 x 4: a
// thread per fine cell of 4 whole rows of up to 256 cells (one coarse row, so no
// x halo is read from a neighbouring block), marching kRR4Planes planes with x and
// the z coupling in registers, as k_true_residual64; consecutive blocks of one XCD
// take y-adjacent coarse rows, so the rows above and below a block come through
// that XCD's L2. Sum order: the 4 planes of a coarse plane per thread, then the 4
// x-lanes of a coarse cell in pairs (shuffles), then the 4 rows in pairs (LDS).
// Needs ni, nj, nk divisible by 4.
constexpr int kRR4Planes = 16;
constexpr int kRR4X = 256;
static __global__ __launch_bounds__(kRR4X * 4) void k_residual_restrict4(Op<float> op, const double* __restrict__ d64,
                                                                   const double* __restrict__ b64,
                                                                   const float* __restrict__ b32,
                                                                   const double* __restrict__ x,
                                                                   float* __restrict__ crhs,
                                                                   float* __restrict__ csolid,
                                                                   float* __restrict__ cliquid) {
    __shared__ float sr[4][kRR4X / 4], ss[4][kRR4X / 4];
    const int ni = op.ni, nj = op.nj, nk = op.nk, sl = ni * nj;
    const int cx = ni / 4, cy = nj / 4;
    const int tx = threadIdx.x, ty = threadIdx.y;
    // block L runs on XCD L % 8: consecutive logical blocks (x fastest, then the
    // coarse row, then the z chunk) on one XCD
    const int nbx = gridDim.x, nb = nbx * gridDim.y * gridDim.z, L = blockIdx.x + nbx * (blockIdx.y + gridDim.y * blockIdx.z);
    const int xcd = L % 8, lg = xcd * (nb / 8) + min(xcd, nb % 8) + L / 8;
    const int bx = lg % nbx, by = (lg / nbx) % gridDim.y, bz = lg / (nbx * gridDim.y);
    const int i = bx * kRR4X + tx, j = by * 4 + ty;
    const int kb = bz * kRR4Planes, ke = min(kb + kRR4Planes, nk);
    const bool in = i < ni;
    int c = (kb * nj + j) * ni + i;
    double xm = (in && kb > 0) ? x[c - sl] : 0.0, x0 = in ? x[c] : 0.0;
    float azm = (in && kb > 0) ? op.AZ(c - sl) : 0.f;
    float acc = 0.f, cnt = 0.f;
    for (int k = kb; k < ke; ++k, c += sl) {
        if (in) {
            const double xp = k + 1 < nk ? x[c + sl] : 0.0;
            const float az0 = k + 1 < nk ? op.AZ(c) : 0.f;
            double y = (d64 ? d64[c] : (double)op.D(c)) * x0;
            if (k > 0) y += (double)azm * xm;
            if (j > 0) y += (double)op.AY(c - ni) * x[c - ni];
            if (i > 0) y += (double)op.AX(c - 1) * x[c - 1];
            if (i + 1 < ni) y += (double)op.AX(c) * x[c + 1];
            if (j + 1 < nj) y += (double)op.AY(c) * x[c + ni];
            if (k + 1 < nk) y += (double)az0 * xp;
            acc += (float)((b64 ? b64[c] : (double)b32[c]) - y);
            cnt += op.D(c) == 0.f ? 1.f : 0.f;
            xm = x0;
            x0 = xp;
            azm = az0;
        }
        if ((k & 3) == 3) {  // the last plane of a coarse plane (uniform over the block)
            acc += __shfl_xor(acc, 1, 64);
            acc += __shfl_xor(acc, 2, 64);
            cnt += __shfl_xor(cnt, 1, 64);
            cnt += __shfl_xor(cnt, 2, 64);
            if ((tx & 3) == 0) {
                sr[ty][tx >> 2] = acc;
                ss[ty][tx >> 2] = cnt;
            }
            __syncthreads();
            const int ic = bx * (kRR4X / 4) + tx;
            if (ty == 0 && tx < kRR4X / 4 && ic < cx) {
                const float inv = 1.0f / 64.0f;
                const float r = ((sr[0][tx] + sr[1][tx]) + (sr[2][tx] + sr[3][tx])) * inv;
                const float so = ((ss[0][tx] + ss[1][tx]) + (ss[2][tx] + ss[3][tx])) * inv;
                const size_t cc = ((size_t)(k >> 2) * cy + by) * cx + ic;
                crhs[cc] = r;
                csolid[cc] = so;
                cliquid[cc] = 1.f - so;
            }
            __syncthreads();
            acc = 0.f;
            cnt = 0.f;
        }
    }
}

// Energy of an FP32 initial guess before its FP64 residual: partial sums of
// x.(2b - A x), A x from the FP32 level-0 operator accumulated in FP64, so that
// E = -sum / 2 as from k_true_residual64's energy sums (x.(b + r) with r = b - A x). Launched as
// k_true_residual64.
static __global__ __launch_bounds__(kBlock) void k_energy_f32(Op<float> op, const double* __restrict__ b64,
                                                       const float* __restrict__ b32, const float* __restrict__ x,
                                                       double* __restrict__ partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int kb = blockIdx.z * kTrueResKZ;
    const int ni = op.ni, nj = op.nj, nk = op.nk, sl = ni * nj;
    double acc = 0;
    if (i < ni && j < nj && kb < nk) {
        for (int k = kb; k < min(kb + kTrueResKZ, nk); ++k) {
            const int c = (k * nj + j) * ni + i;
            const double xc = x[c];
            double y = (double)op.D(c) * xc;
            if (k > 0) y += (double)op.AZ(c - sl) * x[c - sl];
            if (j > 0) y += (double)op.AY(c - ni) * x[c - ni];
            if (i > 0) y += (double)op.AX(c - 1) * x[c - 1];
            if (i + 1 < ni) y += (double)op.AX(c) * x[c + 1];
            if (j + 1 < nj) y += (double)op.AY(c) * x[c + ni];
            if (k + 1 < nk) y += (double)op.AZ(c) * x[c + sl];
            acc += xc * (2.0 * (b64 ? b64[c] : (double)b32[c]) - y);
        }
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partial[(blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = t;
}

// x64 += alpha s; r -= alpha q; partial max|r| (mixed mode, FP64 iterate)
static __global__ __launch_bounds__(kBlock) void k_update_x64r(unsigned n, const Scalars* __restrict__ sc,
                                                        double* __restrict__ x, const float* __restrict__ s,
                                                        float* __restrict__ r, const float* __restrict__ q,
                                                        double* __restrict__ partial) {
    const double alpha = sc->alpha;
    float m = 0.f;
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) {
        x[c] += alpha * (double)s[c];
        const float rv = (float)((double)r[c] - alpha * (double)q[c]);
        r[c] = rv;
        m = fmaxf(m, fabsf(rv));
    }
    float t = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = (double)t;
}

static __global__ __launch_bounds__(kBlock) void k_absmax64(unsigned n, const double* __restrict__ a,
                                                     double* __restrict__ partial) {
    double m = 0;
    if (n % 2 == 0 && ((uintptr_t)a & 15) == 0) {
        // two cells per load (a maximum does not depend on the order)
        const double2* a2 = reinterpret_cast<const double2*>(a);
#pragma unroll 2
        for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n / 2; c += gridDim.x * kBlock) {
            const double2 v = a2[c];
            m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
        }
    } else {
        for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) m = fmax(m, fabs(a[c]));
    }
    double t = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

static __global__ __launch_bounds__(kBlock) void k_cast_f64_f32(unsigned n, const double* __restrict__ a,
                                                         float* __restrict__ o) {
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) o[c] = (float)a[c];
}
// The same two cells per thread (16-byte loads, 8-byte stores; n even, a 16-byte and
// o 8-byte aligned); the grid covers n / 2 threads.
static __global__ __launch_bounds__(kBlock) void k_cast_f64_f32_x2(unsigned n, const double* __restrict__ a,
                                                            float* __restrict__ o) {
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n / 2) {
        const double2 v = reinterpret_cast<const double2*>(a)[c];
        reinterpret_cast<float2*>(o)[c] = make_float2((float)v.x, (float)v.y);
    }
}
static __global__ __launch_bounds__(kBlock) void k_cast_f32_f64(unsigned n, const float* __restrict__ a,
                                                         double* __restrict__ o) {
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) o[c] = (double)a[c];
}
// Energy-optimal initial guess from the pressure history (the projection onto
// previous solutions of P. F. Fischer, "Projection techniques for iterative
// solution of Ax = b with successive right-hand sides", CMAME 163, 1998): the
// x0 in span{p1, p2, p3} that minimises x'Ax/2 - b'x, from the Gram matrix
// G = P'AP (kept incrementally: each history solve's final true residual adds its
// row) and f = P'b (one streaming pass per guess).
constexpr int kHistSums = 9;

// Commits a solve's Gram row: sums the partials of k_true_residual64 in a fixed order
// into G (3 x 3 over history slots, FP64): G[s][s] = x.Ax, G[s][a] = G[a][s] = p_a.Ax,
// likewise for b (a, b < 0: absent). One workgroup.
static __global__ __launch_bounds__(kBlock) void k_hist_gram_commit(const double* __restrict__ part, int nb, int s,
                                                                    int a, int b, double* __restrict__ G) {
    double v[3];
    for (int q = 0; q < 3; ++q) {
        double t = 0;
        for (int i = threadIdx.x; i < nb; i += kBlock) t += part[q * nb + i];
        v[q] = block_sum(t);
    }
    if (threadIdx.x != 0) return;
    G[3 * s + s] = v[0];
    if (a >= 0) G[3 * s + a] = G[3 * a + s] = v[1];
    if (b >= 0) G[3 * s + b] = G[3 * b + s] = v[2];
}

// Partial sums of f_i = p_i.b over the history entries (M of them), FP64; grid-stride,
// partials at partial[i * gridDim.x + block].
template <int M>
static __global__ __launch_bounds__(kBlock) void k_hist_rhs_dots(unsigned n, const double* __restrict__ b64,
                                                                 const float* __restrict__ b32,
                                                                 const double* __restrict__ p1,
                                                                 const double* __restrict__ p2,
                                                                 const double* __restrict__ p3,
                                                                 double* __restrict__ partial) {
    double f[3] = {0, 0, 0};
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) {
        const double bc = b64 ? b64[c] : (double)b32[c];
        f[0] += p1[c] * bc;
        if (M >= 2) f[1] += p2[c] * bc;
        if (M >= 3) f[2] += p3[c] * bc;
    }
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const double t = block_sum(f[q]);
        if (threadIdx.x == 0) partial[q * gridDim.x + blockIdx.x] = t;
    }
}

// k_hist_rhs_dots two cells per thread and step (16-byte loads; arrays 16-byte
// aligned, the FP32 right-hand side 8-byte aligned; n even).
template <int M>
static __global__ __launch_bounds__(kBlock) void k_hist_rhs_dots2(unsigned n, const double* __restrict__ b64,
                                                                  const float* __restrict__ b32,
                                                                  const double* __restrict__ p1,
                                                                  const double* __restrict__ p2,
                                                                  const double* __restrict__ p3,
                                                                  double* __restrict__ partial) {
    double f[3] = {0, 0, 0};
    const unsigned n2 = n / 2;
#pragma unroll 2
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n2; c += gridDim.x * kBlock) {
        double2 bc;
        if (b64) {
            bc = reinterpret_cast<const double2*>(b64)[c];
        } else {
            const float2 t = reinterpret_cast<const float2*>(b32)[c];
            bc = make_double2((double)t.x, (double)t.y);
        }
        const double2 a = reinterpret_cast<const double2*>(p1)[c];
        f[0] += a.x * bc.x;
        f[0] += a.y * bc.y;
        if (M >= 2) {
            const double2 q = reinterpret_cast<const double2*>(p2)[c];
            f[1] += q.x * bc.x;
            f[1] += q.y * bc.y;
        }
        if (M >= 3) {
            const double2 q = reinterpret_cast<const double2*>(p3)[c];
            f[2] += q.x * bc.x;
            f[2] += q.y * bc.y;
        }
    }
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const double t = block_sum(f[q]);
        if (threadIdx.x == 0) partial[q * gridDim.x + blockIdx.x] = t;
    }
}

// One workgroup: f from its partials (fixed order), G over the slots (sa, sb, sc) of
// p1, p2, p3 from the Gram cache; both taken to the difference basis v0 = p1,
// v1 = p1 - p2, v2 = p1 - 2 p2 + p3 (exact FP64 combinations: the last solutions are
// nearly parallel, their differences keep the system well conditioned); G c = f
// solved by symmetric elimination (a basis vector whose pivot falls below 1e-13 of
// the first diagonal is dropped: coefficient 0); coef[0..2] = the weights of p1, p2,
// p3 in x0 = c0 v0 + c1 v1 + c2 v2.
// M is a template parameter so that the small arrays below stay in registers (a
// runtime M indexes them dynamically, i.e. through scratch memory).
template <int M>
static __global__ __launch_bounds__(kBlock) void k_hist_solve(const double* __restrict__ partial, int nb,
                                                              const double* __restrict__ Gc, int sa, int sb, int sc,
                                                              double* __restrict__ coef) {
    __shared__ double fs[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double s = 0;
        if (q < M)
            for (int b = threadIdx.x; b < nb; b += kBlock) s += partial[q * nb + b];
        s = block_sum(s);
        if (threadIdx.x == 0) fs[q] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int sl[3] = {sa, sb, sc};
    double Gp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int i = 0; i < M; ++i)
    #pragma unroll
    for (int j = 0; j < M; ++j) Gp[i][j] = Gc[3 * sl[i] + sl[j]];
    // T: columns are v0, v1, v2 in the p basis
    const double T[3][3] = {{1, 1, 1}, {0, -1, -2}, {0, 0, 1}};
    double G[3][3], f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        f[k] = 0;
    #pragma unroll
    for (int i = 0; i < M; ++i) f[k] += T[i][k] * fs[i];
    #pragma unroll
    for (int l = 0; l < 3; ++l) {
            double t = 0;
        #pragma unroll
    for (int i = 0; i < M; ++i)
            #pragma unroll
    for (int j = 0; j < M; ++j) t += T[i][k] * Gp[i][j] * T[j][l];
            G[k][l] = t;
        }
    }
    double c[3] = {0, 0, 0};
    bool keep[3] = {false, false, false};
    const double ref = G[0][0];
    if (ref > 0) {
    #pragma unroll
    for (int k = 0; k < M; ++k) {
            if (!(G[k][k] > 1e-13 * ref)) continue;
            keep[k] = true;
        #pragma unroll
    for (int r = k + 1; r < M; ++r) {
                const double l = G[r][k] / G[k][k];
            #pragma unroll
    for (int q = k; q < M; ++q) G[r][q] -= l * G[k][q];
                f[r] -= l * f[k];
            }
        }
    #pragma unroll
    for (int k = M - 1; k >= 0; --k) {
            if (!keep[k]) continue;
            double t = f[k];
        #pragma unroll
    for (int q = k + 1; q < M; ++q) t -= G[k][q] * c[q];
            c[k] = t / G[k][k];
        }
    }
    coef[0] = c[0] + c[1] + c[2];
    coef[1] = -(c[1] + 2.0 * c[2]);
    coef[2] = c[2];
}

// g = coef[0] p1 + coef[1] p2 + coef[2] p3 (g may alias p3: element-wise).
static __global__ __launch_bounds__(kBlock) void k_hist_combine(unsigned n, const double* __restrict__ coef,
                                                                const double* p1, const double* p2, const double* p3,
                                                                double* g) {
    const double a0 = coef[0], a1 = coef[1], a2 = coef[2];
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) {
        double v = a0 * p1[c];
        if (p2) v += a1 * p2[c];
        if (p3) v += a2 * p3[c];
        g[c] = v;
    }
}

// Extrapolated initial guess from the pressure history: g = 2 p1 - p2 (p2 null: g = p1).
static __global__ __launch_bounds__(kBlock) void k_extrapolate64(unsigned n, const double* __restrict__ p1,
                                                                 const double* __restrict__ p2, double* __restrict__ g) {
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock)
        g[c] = p2 ? 2.0 * p1[c] - p2[c] : p1[c];
}
// g += (double)d
static __global__ __launch_bounds__(kBlock) void k_add_f32_f64(unsigned n, const float* __restrict__ d,
                                                               double* __restrict__ g) {
    for (unsigned c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) g[c] += (double)d[c];
}
static __global__ __launch_bounds__(kBlock) void k_fill64(unsigned n, double* __restrict__ a, double v) {
    unsigned c = blockIdx.x * kBlock + threadIdx.x;
    if (c < n) a[c] = v;
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
__device__ inline void dot_body(unsigned n, const float* __restrict__ a, const float* __restrict__ b,
                                double* __restrict__ partial, Blk B) {
    Acc acc = 0;
    for (unsigned c = B.x * kBlock + threadIdx.x; c < n; c += B.nx * kBlock) acc += (Acc)a[c] * (Acc)b[c];
    Acc t = block_sum(acc);
    if (threadIdx.x == 0) partial[B.x] = (double)t;
}
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_dot(unsigned n, const float* __restrict__ a, const float* __restrict__ b,
                                                double* __restrict__ partial) {
    dot_body<Acc>(n, a, b, partial, this_blk());
}

__device__ inline void absmax_body(unsigned n, const float* __restrict__ a, double* __restrict__ partial, Blk B) {
    float m = 0.f;
    for (unsigned c = B.x * kBlock + threadIdx.x; c < n; c += B.nx * kBlock) m = fmaxf(m, fabsf(a[c]));
    float t = block_max(m);
    if (threadIdx.x == 0) partial[B.x] = (double)t;
}
static __global__ __launch_bounds__(kBlock) void k_absmax(unsigned n, const float* __restrict__ a, double* __restrict__ partial) {
    absmax_body(n, a, partial, this_blk());
}

// x += alpha s; r -= alpha q; partial max|r| (S/AlgebraicMultigrid.h:259-262)
template <typename Acc>
__device__ inline void update_xr_body(unsigned n, const Scalars* __restrict__ sc, float* __restrict__ x,
                                      const float* __restrict__ s, float* __restrict__ r, const float* __restrict__ q,
                                      double* __restrict__ partial, Blk B) {
    const Acc alpha = (Acc)sc->alpha;
    float m = 0.f;
    for (unsigned c = B.x * kBlock + threadIdx.x; c < n; c += B.nx * kBlock) {
        x[c] = (float)((Acc)x[c] + alpha * (Acc)s[c]);
        const float rv = (float)((Acc)r[c] - alpha * (Acc)q[c]);
        r[c] = rv;
        m = fmaxf(m, fabsf(rv));
    }
    float t = block_max(m);
    if (threadIdx.x == 0) partial[B.x] = (double)t;
}
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_update_xr(unsigned n, const Scalars* __restrict__ sc, float* __restrict__ x,
                                                      const float* __restrict__ s, float* __restrict__ r,
                                                      const float* __restrict__ q, double* __restrict__ partial) {
    update_xr_body<Acc>(n, sc, x, s, r, q, partial, this_blk());
}

// s = z + beta s (S/AlgebraicMultigrid.h:283)
template <typename Acc>
__device__ inline void update_s_body(unsigned n, const Scalars* __restrict__ sc, float* __restrict__ s,
                                     const float* __restrict__ z, Blk B) {
    const Acc beta = (Acc)sc->beta;
    for (unsigned c = B.x * kBlock + threadIdx.x; c < n; c += B.nx * kBlock) s[c] = (float)((Acc)z[c] + beta * (Acc)s[c]);
}
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_update_s(unsigned n, const Scalars* __restrict__ sc, float* __restrict__ s,
                                                     const float* __restrict__ z) {
    update_s_body<Acc>(n, sc, s, z, this_blk());
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
__device__ inline void finalize_body(Scalars* sc, const double* __restrict__ partial, int np, int what, double tol) {
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
    } else if (what == 6) {
        sc->energy = -0.5 * v;
    } else {
        sc->true_res = v;
    }
}
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_finalize(Scalars* sc, const double* __restrict__ partial, int np, int what,
                                                     double tol) {
    finalize_body<Acc>(sc, partial, np, what, tol);
}

__device__ inline void fill_body(unsigned n, float* __restrict__ a, float v, Blk B) {
    unsigned c = B.x * kBlock + threadIdx.x;
    if (c < n) a[c] = v;
}
static __global__ __launch_bounds__(kBlock) void k_fill(unsigned n, float* __restrict__ a, float v) {
    fill_body(n, a, v, this_blk());
}
__device__ inline void copy_body(unsigned n, float* __restrict__ dst, const float* __restrict__ src, Blk B) {
    unsigned c = B.x * kBlock + threadIdx.x;
    if (c < n) dst[c] = src[c];
}
static __global__ __launch_bounds__(kBlock) void k_copy(unsigned n, float* __restrict__ dst, const float* __restrict__ src) {
    copy_body(n, dst, src, this_blk());
}

}  // namespace vflip
