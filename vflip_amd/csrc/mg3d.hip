// Geometric multigrid for the constant-coefficient 7-point Poisson problem on a
// uniform cell grid: MultiGridSolver3D<T> of the reference
// (S/Multigrid3D.h:124-137, 840-1436) with T = float or double.
//
// Operator (S/Multigrid3D.h:1130-1186): (A x)(c) = sum of the six neighbour
// values - 6 x(c), where a neighbour index past the grid is clamped to the cell
// itself (Neumann) or, with the Dirichlet flag, reads 0. The smoother
// (840-1129) is red-black Gauss-Seidel with weight omega, parity-1 cells first:
//   x = (1 - omega) x + (omega * (sum - b)) / 6,   sum/b in T, the blend in double.
// Levels halve every extent (floor, at least 1) until the grid has at most 512
// cells (124-137); restriction is 0.5 * the sum of the 8 children of the
// residual (1193-1245), prolongation adds the parent's value (1298-1326), and
// the bottom level runs 100 sweeps with omega = 1 (1187-1191). A V-cycle
// (1328-1392) smooths 5 sweeps down and 5 up; full multigrid (1394-1436)
// restricts the residual to the bottom, then prolongs and V-cycles every level
// and finishes with one more V-cycle on the finest.
//
// MI355X layout. One fused launch per smoothing sweep: a workgroup owns a
// 64 x 8 column of cells and marches up a z chunk with three planes of the old
// values (and a 2-cell ring) in LDS. At plane p it updates the parity-1 cells
// of plane p over the tile plus a 1-cell ring, then the parity-0 cells of
// plane p-1, and writes plane p-1 whole. Old values come from one buffer and
// new ones go to another (x and x_new of the reference), so the ring's
// parity-1 cells, recomputed by both neighbouring workgroups, are the same
// numbers and every cell is written by exactly one workgroup: HBM traffic is
// one read of x and b and one write of x per sweep. Zero initial values and
// the prolongation of the coarser level's correction are folded into the
// first sweep's loads, the residual into the restriction, and the bottom
// level (<= 512 cells) runs all its sweeps in one workgroup in LDS. Results
// are bit-identical with the reference's arithmetic: the file is compiled
// without FMA contraction and every sum keeps the reference's order.
#pragma clang fp contract(off)

#include <algorithm>
#include <vector>

#include "common.h"

namespace vflip {
namespace {

constexpr int kTX = 64, kTY = 8;              // output tile of a sweep workgroup
constexpr int kLX = kTX + 4, kLY = kTY + 4;   // old values with a 2-cell ring
constexpr int kBX = kTX + 2, kBY = kTY + 2;   // new parity-1 values with a 1-cell ring
constexpr int kSweepThreads = 256;
constexpr int kLoadsPerPlane = (kLX * kLY + kSweepThreads - 1) / kSweepThreads;
constexpr int kBlackCells = kBY * (kBX / 2);  // parity-1 cells of the tile + ring on one plane
constexpr int kBlackPerThread = (kBlackCells + kSweepThreads - 1) / kSweepThreads;
constexpr int kRowsPerThread = kTX * kTY / kSweepThreads;
static_assert(kTX == 64 && kTX * kTY % kSweepThreads == 0, "a wave spans one tile row");
constexpr int kBottomCells = 1024;            // one workgroup, one cell per thread
constexpr int kMaxLevels = 32;
constexpr int kCoarsest = 512;                // m_ComputeLevels' stopping size

// ((l + r) + (t + d)) + (f + bk) in T: the reference's accumulation order.
template <typename T> __device__ __forceinline__ T nsum(T lv, T rv, T tv, T dv, T fv, T bv) {
    T rr = lv + rv;
    rr += tv + dv;
    rr += fv + bv;
    return rr;
}

// (1 - omega) * x + omega * rr / 6.0 with double omega (S/Multigrid3D.h:1070,1119).
template <typename T> __device__ __forceinline__ T relax(T self, T rr, double omega) {
    return (T)((1.0 - omega) * (double)self + (omega * (double)rr) / 6.0);
}

// Old value of a cell for the first sweep of a level.
//   MODE 0: xin             MODE 1: 0 + parent      MODE 2: xin + parent      MODE 3: 0
// The parent of (i,j,k) is (i/2, j/2, k/2) on the coarser grid; a parent past
// that grid (odd fine extents) contributes 0.
template <typename T, int MODE>
__device__ __forceinline__ T old_value(const T* __restrict__ xin, const T* __restrict__ xc, int i, int j, int k,
                                       int nx, int ny, int cx, int cy, int cz) {
    T v = T(0);
    if constexpr (MODE == 0 || MODE == 2) v = xin[lin3(i, j, k, nx, ny)];
    if constexpr (MODE == 1 || MODE == 2) {
        const int I = i >> 1, J = j >> 1, K = k >> 1;
        const T pv = (I < cx && J < cy && K < cz) ? xc[lin3(I, J, K, cx, cy)] : T(0);
        v += pv;
    }
    return v;
}

// Blocks on one XCD (block L runs on XCD L % 8) get consecutive logical tiles.
__device__ __forceinline__ int xcd_tile() {
    const int nb = gridDim.x, L = blockIdx.x, xcd = L % 8;
    return xcd * (nb / 8) + min(xcd, nb % 8) + L / 8;
}

// One red-black sweep, old values from (xin, xc) per MODE, result to xout.
// Loads run one plane ahead of the arithmetic: the old values of plane p+2 and
// the right-hand sides of the next plane's cells are in flight while plane p
// is updated.
template <typename T, int MODE>
__global__ __launch_bounds__(kSweepThreads) void k_mg3d_sweep(int nx, int ny, int nz, int dirichlet, double omega,
                                                              const T* __restrict__ b, const T* __restrict__ xin,
                                                              const T* __restrict__ xc, int cx, int cy, int cz,
                                                              T* __restrict__ xout, int tiles_x, int tiles_y,
                                                              int zchunk) {
    __shared__ T X[3][kLY][kLX];
    __shared__ T B[3][kBY][kBX];
    const int lg = xcd_tile();
    const int tx = lg % tiles_x, ty = (lg / tiles_x) % tiles_y, tz = lg / (tiles_x * tiles_y);
    const int i0 = tx * kTX, j0 = ty * kTY, k0 = tz * zchunk, k1 = min(k0 + zchunk, nz);
    const int tid = threadIdx.x;
    const T zero = T(0);
    auto edge = [&](T self) { return dirichlet ? zero : self; };

    // the loads of one plane of old values (2-cell ring, indexes clamped into the grid)
    auto fetch = [&](int q, T (&v)[kLoadsPerPlane]) {
        const int k = min(max(q, 0), nz - 1);
#pragma unroll
        for (int u = 0; u < kLoadsPerPlane; ++u) {
            const int idx = tid + u * kSweepThreads;
            if (idx < kLX * kLY) {
                const int r = idx / kLX, c = idx - r * kLX;
                const int i = min(max(i0 - 2 + c, 0), nx - 1), j = min(max(j0 - 2 + r, 0), ny - 1);
                v[u] = old_value<T, MODE>(xin, xc, i, j, k, nx, ny, cx, cy, cz);
            }
        }
    };
    auto put = [&](int slot, const T (&v)[kLoadsPerPlane]) {
#pragma unroll
        for (int u = 0; u < kLoadsPerPlane; ++u) {
            const int idx = tid + u * kSweepThreads;
            if (idx < kLX * kLY) (&X[slot][0][0])[idx] = v[u];
        }
    };
    auto slot = [](int q) { return (q + 6) % 3; };
    // parity-1 cell u of this thread on plane q (tile + 1-cell ring): ring row r, column c
    auto black_cell = [&](int u, int q, int& r, int& c, int& i, int& j) {
        const int idx = tid + u * kSweepThreads;
        r = idx / (kBX / 2);
        const int cc = idx - r * (kBX / 2);
        j = j0 - 1 + r;
        c = 2 * cc + ((i0 + j + q) & 1);
        i = i0 - 1 + c;
        return idx < kBlackCells && i >= 0 && i < nx && j >= 0 && j < ny && q >= 0 && q < nz;
    };
    // right-hand sides of this thread's parity-1 cells on plane q
    auto fetch_bb = [&](int q, T (&v)[kBlackPerThread]) {
#pragma unroll
        for (int u = 0; u < kBlackPerThread; ++u) {
            int r, c, i, j;
            v[u] = black_cell(u, q, r, c, i, j) ? b[lin3(i, j, q, nx, ny)] : zero;
        }
    };
    // right-hand sides of this thread's tile cells on plane q (parity-0 ones are used)
    auto fetch_br = [&](int q, T (&v)[kRowsPerThread]) {
#pragma unroll
        for (int u = 0; u < kRowsPerThread; ++u) {
            const int r = tid / kTX + u * (kSweepThreads / kTX), c = tid % kTX;
            const int i = i0 + c, j = j0 + r;
            v[u] = (i < nx && j < ny && q >= k0 && q < k1 && ((i + j + q) & 1) == 0) ? b[lin3(i, j, q, nx, ny)] : zero;
        }
    };

    T pf[kLoadsPerPlane], bb[kBlackPerThread], br[kRowsPerThread];
    fetch(k0 - 2, pf);
    put(slot(k0 - 2), pf);
    fetch(k0 - 1, pf);
    put(slot(k0 - 1), pf);
    fetch(k0, pf);
    fetch_bb(k0 - 1, bb);
#pragma unroll
    for (int u = 0; u < kRowsPerThread; ++u) br[u] = zero;

    for (int p = k0 - 1; p <= k1; ++p) {
        const int sm1 = slot(p - 1), s0 = slot(p), sp1 = slot(p + 1);
        lds_barrier();  // every read of the slots reused below is done
        put(sp1, pf);
        if (p + 2 <= k1 + 1) fetch(p + 2, pf);
        T bbn[kBlackPerThread], brn[kRowsPerThread];
        fetch_bb(p + 1, bbn);
        fetch_br(p, brn);
        lds_barrier();

        // parity-1 cells of plane p over the tile and a 1-cell ring
#pragma unroll
        for (int u = 0; u < kBlackPerThread; ++u) {
            int r, c, i, j;
            if (black_cell(u, p, r, c, i, j)) {
                const int xr = r + 1, xcl = c + 1;
                const T self = X[s0][xr][xcl];
                const T lv = i == 0 ? edge(self) : X[s0][xr][xcl - 1];
                const T rv = i == nx - 1 ? edge(self) : X[s0][xr][xcl + 1];
                const T tv = j == ny - 1 ? edge(self) : X[s0][xr + 1][xcl];
                const T dv = j == 0 ? edge(self) : X[s0][xr - 1][xcl];
                const T fv = p == 0 ? edge(self) : X[sm1][xr][xcl];
                const T bv = p == nz - 1 ? edge(self) : X[sp1][xr][xcl];
                const T rr = nsum(lv, rv, tv, dv, fv, bv) - bb[u];
                B[s0][r][c] = relax(self, rr, omega);
            }
        }
        lds_barrier();

        // parity-0 cells of plane p-1, then plane p-1 of the tile written whole
        const int k = p - 1;
        if (k >= k0 && k < k1) {
            const int bm2 = slot(p - 2), bm1 = slot(p - 1), b0 = slot(p);
#pragma unroll
            for (int u = 0; u < kRowsPerThread; ++u) {
                const int r = tid / kTX + u * (kSweepThreads / kTX), c = tid % kTX;
                const int i = i0 + c, j = j0 + r;
                if (i < nx && j < ny) {
                    const int br_ = r + 1, bc = c + 1;
                    T out;
                    if (((i + j + k) & 1) == 0) {
                        const T self = X[sm1][r + 2][c + 2];
                        const T lv = i == 0 ? edge(self) : B[bm1][br_][bc - 1];
                        const T rv = i == nx - 1 ? edge(self) : B[bm1][br_][bc + 1];
                        const T tv = j == ny - 1 ? edge(self) : B[bm1][br_ + 1][bc];
                        const T dv = j == 0 ? edge(self) : B[bm1][br_ - 1][bc];
                        const T fv = k == 0 ? edge(self) : B[bm2][br_][bc];
                        const T bv = k == nz - 1 ? edge(self) : B[b0][br_][bc];
                        const T rr = nsum(lv, rv, tv, dv, fv, bv) - br[u];
                        out = relax(self, rr, omega);
                    } else {
                        out = B[bm1][br_][bc];
                    }
                    xout[lin3(i, j, k, nx, ny)] = out;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kBlackPerThread; ++u) bb[u] = bbn[u];
#pragma unroll
        for (int u = 0; u < kRowsPerThread; ++u) br[u] = brn[u];
    }
}

// Residual b - A x at one cell (S/Multigrid3D.h:1130-1186); ZX: x is zero.
template <typename T, bool ZX>
__device__ __forceinline__ T residual_at(const T* __restrict__ b, const T* __restrict__ x, int i, int j, int k, int nx,
                                         int ny, int nz, int dirichlet) {
    const T rhs = b[lin3(i, j, k, nx, ny)];
    if constexpr (ZX) {
        const T z = T(0);
        T rr = nsum(z, z, z, z, z, z);
        rr -= T(6) * z;
        return rhs - rr;
    } else {
        const T zero = T(0);
        const T center = x[lin3(i, j, k, nx, ny)];
        T lv = x[lin3(i == 0 ? i : i - 1, j, k, nx, ny)];
        T rv = x[lin3(i == nx - 1 ? i : i + 1, j, k, nx, ny)];
        T tv = x[lin3(i, j == ny - 1 ? j : j + 1, k, nx, ny)];
        T dv = x[lin3(i, j == 0 ? j : j - 1, k, nx, ny)];
        T fv = x[lin3(i, j, k == 0 ? k : k - 1, nx, ny)];
        T bv = x[lin3(i, j, k == nz - 1 ? k : k + 1, nx, ny)];
        if (dirichlet) {
            if (i == 0) lv = zero;
            if (i == nx - 1) rv = zero;
            if (j == ny - 1) tv = zero;
            if (j == 0) dv = zero;
            if (k == 0) fv = zero;
            if (k == nz - 1) bv = zero;
        }
        T rr = nsum(lv, rv, tv, dv, fv, bv);
        rr -= T(6) * center;
        return rhs - rr;
    }
}

template <typename T>
__global__ void k_mg3d_residual(int nx, int ny, int nz, int dirichlet, const T* __restrict__ b,
                                const T* __restrict__ x, T* __restrict__ res) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (t >= n) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((size_t)nx * ny));
    res[t] = residual_at<T, false>(b, x, i, j, k, nx, ny, nz, dirichlet);
}

// Coarse cell (I,J,K) = 0.5 * sum of its 8 children in the reference's order
// (S/Multigrid3D.h:1219-1238); a child past the fine grid (extent 1) adds 0.
// RES: the children are residuals of (b, x) computed on the fly; otherwise src.
template <typename T, bool RES, bool ZX>
__global__ void k_mg3d_restrict(int nx, int ny, int nz, int dirichlet, const T* __restrict__ src,
                                const T* __restrict__ x, int cx, int cy, int cz, T* __restrict__ dst) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)cx * cy * cz;
    if (t >= n) return;
    const int I = (int)(t % cx), J = (int)((t / cx) % cy), K = (int)(t / ((size_t)cx * cy));
    T r = T(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
        T v = T(0);
        if (i < nx && j < ny && k < nz) {
            if constexpr (RES) v = residual_at<T, ZX>(src, x, i, j, k, nx, ny, nz, dirichlet);
            else v = src[lin3(i, j, k, nx, ny)];
        }
        r += v;
    }
    dst[t] = (T)((double)r * 0.5);
}

// dst(i,j,k) += src(i/2, j/2, k/2) (S/Multigrid3D.h:1298-1326); parents past
// the coarse grid add 0.
template <typename T>
__global__ void k_mg3d_prolongate(int nx, int ny, int nz, const T* __restrict__ src, int cx, int cy, int cz,
                                  T* __restrict__ dst) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    if (t >= n) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((size_t)nx * ny));
    const int I = i >> 1, J = j >> 1, K = k >> 1;
    const T pv = (I < cx && J < cy && K < cz) ? src[lin3(I, J, K, cx, cy)] : T(0);
    dst[t] += pv;
}

// `iters` in-place red-black sweeps of a grid of at most kBottomCells cells in
// one workgroup (the bottom solve, S/Multigrid3D.h:1187-1191). ZX: start from 0.
template <typename T, bool ZX>
__global__ __launch_bounds__(kBottomCells) void k_mg3d_bottom(int nx, int ny, int nz, int dirichlet, double omega,
                                                              int iters, const T* __restrict__ b, T* __restrict__ x) {
    __shared__ T xs[kBottomCells];
    const int n = nx * ny * nz, t = threadIdx.x;
    const bool on = t < n;
    const int i = on ? t % nx : 0, j = on ? (t / nx) % ny : 0, k = on ? t / (nx * ny) : 0;
    const T zero = T(0);
    T bt = zero;
    if (on) {
        xs[t] = ZX ? zero : x[t];
        bt = b[t];
    }
    __syncthreads();
    const int parity = (i + j + k) & 1;
    const int sx = 1, sy = nx, sz = nx * ny;
    for (int it = 0; it < iters; ++it) {
        for (int color = 1; color >= 0; --color) {
            if (on && parity == color) {
                const T self = xs[t];
                T lv = i == 0 ? self : xs[t - sx];
                T rv = i == nx - 1 ? self : xs[t + sx];
                T tv = j == ny - 1 ? self : xs[t + sy];
                T dv = j == 0 ? self : xs[t - sy];
                T fv = k == 0 ? self : xs[t - sz];
                T bv = k == nz - 1 ? self : xs[t + sz];
                if (dirichlet) {
                    if (i == 0) lv = zero;
                    if (i == nx - 1) rv = zero;
                    if (j == ny - 1) tv = zero;
                    if (j == 0) dv = zero;
                    if (k == 0) fv = zero;
                    if (k == nz - 1) bv = zero;
                }
                const T rr = nsum(lv, rv, tv, dv, fv, bv) - bt;
                xs[t] = relax(self, rr, omega);
            }
            __syncthreads();
        }
    }
    if (on) x[t] = xs[t];
}

struct Grid {
    int n[3];
    size_t cells() const { return (size_t)n[0] * n[1] * n[2]; }
};

Grid coarser(const Grid& g) {
    Grid c;
    for (int a = 0; a < 3; ++a) c.n[a] = std::max(g.n[a] / 2, 1);
    return c;
}

// Level extents of MultiGrid::m_ComputeLevels / m_AssignMemoryEachLevel (S/Multigrid3D.h:100-137).
std::vector<Grid> level_grids(int nx, int ny, int nz) {
    std::vector<Grid> g{{{nx, ny, nz}}};
    while (g.back().cells() > (size_t)kCoarsest) g.push_back(coarser(g.back()));
    return g;
}

template <typename T> struct Ops {
    hipStream_t s;
    int dirichlet;

    // First-sweep modes: which of (xin, xc) feed the old values.
    static int mode_of(const T* xin, const T* xc) { return xin ? (xc ? 2 : 0) : (xc ? 1 : 3); }

    int sweep(const Grid& g, const T* b, const T* xin, const T* xc, const Grid& cg, T* xout, double omega) {
        const int tiles_x = (g.n[0] + kTX - 1) / kTX, tiles_y = (g.n[1] + kTY - 1) / kTY;
        const long txy = (long)tiles_x * tiles_y;
        // about 2048 workgroups; a chunk of at least 2 planes
        int zchunk = (int)std::min<long>(g.n[2], std::max<long>(2, (g.n[2] * txy + 2047) / 2048));
        const int tiles_z = (g.n[2] + zchunk - 1) / zchunk;
        const unsigned blocks = (unsigned)(txy * tiles_z);
        const int m = mode_of(xin, xc);
#define VF_SWEEP(M)                                                                                                  \
    k_mg3d_sweep<T, M><<<blocks, kSweepThreads, 0, s>>>(g.n[0], g.n[1], g.n[2], dirichlet, omega, b, xin, xc,         \
                                                        cg.n[0], cg.n[1], cg.n[2], xout, tiles_x, tiles_y, zchunk)
        if (m == 0) VF_SWEEP(0);
        else if (m == 1) VF_SWEEP(1);
        else if (m == 2) VF_SWEEP(2);
        else VF_SWEEP(3);
#undef VF_SWEEP
        VF_CHECK_LAUNCH();
        return 0;
    }

    // `sweeps` sweeps ping-ponging between bufA and bufB; the first reads
    // (xin, xc) per mode_of. *out is the buffer holding the result.
    int smooth(const Grid& g, const T* b, const T* xin, const T* xc, const Grid& cg, T* bufA, T* bufB, int sweeps,
               double omega, T** out) {
        const T* cur = xin;
        const T* pro = xc;
        T* res = const_cast<T*>(xin);
        for (int q = 0; q < sweeps; ++q) {
            T* dst = (cur == bufA) ? bufB : bufA;
            int rc = sweep(g, b, cur, pro, cg, dst, omega);
            if (rc) return rc;
            cur = dst;
            res = dst;
            pro = nullptr;
        }
        *out = res;
        return 0;
    }

    int bottom(const Grid& g, const T* b, T* x, bool zero_start, int iters, double omega) {
        if (zero_start)
            k_mg3d_bottom<T, true><<<1, kBottomCells, 0, s>>>(g.n[0], g.n[1], g.n[2], dirichlet, omega, iters, b, x);
        else
            k_mg3d_bottom<T, false><<<1, kBottomCells, 0, s>>>(g.n[0], g.n[1], g.n[2], dirichlet, omega, iters, b, x);
        VF_CHECK_LAUNCH();
        return 0;
    }

    int resid_restrict(const Grid& g, const T* b, const T* x, const Grid& cg, T* bc) {
        if (x)
            k_mg3d_restrict<T, true, false><<<grid_for(cg.cells()), kBlock, 0, s>>>(g.n[0], g.n[1], g.n[2], dirichlet,
                                                                                    b, x, cg.n[0], cg.n[1], cg.n[2], bc);
        else
            k_mg3d_restrict<T, true, true><<<grid_for(cg.cells()), kBlock, 0, s>>>(g.n[0], g.n[1], g.n[2], dirichlet,
                                                                                   b, x, cg.n[0], cg.n[1], cg.n[2], bc);
        VF_CHECK_LAUNCH();
        return 0;
    }
};

}  // namespace
}  // namespace vflip

using namespace vflip;

struct vflip_mg3d {
    int device = 0;
    int type = VFLIP_MG3D_FLOAT;
    int dirichlet = 0;
    std::vector<Grid> grids;
    std::vector<void*> x, t, b;  // per level; x[0], b[0] are the caller's
    std::vector<void*> cur;      // buffer holding each level's current values
    void* pool = nullptr;
};

namespace {

template <typename T> int vcycle_from(vflip_mg3d* M, int L, const T* xin, const T* xc, hipStream_t s) {
    Ops<T> op{s, M->dirichlet};
    const int nl = (int)M->grids.size();
    auto X = [&](int l) { return (T*)M->x[l]; };
    auto Tt = [&](int l) { return (T*)M->t[l]; };
    auto Bv = [&](int l) { return (const T*)M->b[l]; };
    T* cur = nullptr;
    // down: smooth, then the restricted residual becomes the next level's b
    for (int l = L; l < nl - 1; ++l) {
        const T* in = l == L ? xin : nullptr;
        const T* pro = l == L ? xc : nullptr;
        int rc = op.smooth(M->grids[l], Bv(l), in, pro, M->grids[l + 1], X(l), Tt(l), 5, 1.0, &cur);
        if (rc) return rc;
        M->cur[l] = cur;
        rc = op.resid_restrict(M->grids[l], Bv(l), cur, M->grids[l + 1], (T*)M->b[l + 1]);
        if (rc) return rc;
    }
    int rc = op.bottom(M->grids[nl - 1], Bv(nl - 1), X(nl - 1), true, 100, 1.0);
    if (rc) return rc;
    M->cur[nl - 1] = X(nl - 1);
    // up: prolongation folded into the first post-smoothing sweep
    for (int l = nl - 2; l >= L; --l) {
        rc = op.smooth(M->grids[l], Bv(l), (const T*)M->cur[l], (const T*)M->cur[l + 1], M->grids[l + 1], X(l), Tt(l),
                       5, 1.0, &cur);
        if (rc) return rc;
        M->cur[l] = cur;
    }
    return 0;
}

template <typename T> int vcycle_top(vflip_mg3d* M, T* x, const T* b, hipStream_t s) {
    M->x[0] = x;
    M->b[0] = const_cast<T*>(b);
    if (M->grids.size() == 1) {
        Ops<T> op{s, M->dirichlet};
        return op.bottom(M->grids[0], b, x, false, 100, 1.0);
    }
    int rc = vcycle_from<T>(M, 0, x, nullptr, s);
    if (rc) return rc;
    if (M->cur[0] != x) VF_HIP(hipMemcpyAsync(x, M->cur[0], M->grids[0].cells() * sizeof(T), hipMemcpyDeviceToDevice, s));
    return 0;
}

template <typename T> int fmg_top(vflip_mg3d* M, T* x, const T* b, hipStream_t s) {
    M->x[0] = x;
    M->b[0] = const_cast<T*>(b);
    const int nl = (int)M->grids.size();
    Ops<T> op{s, M->dirichlet};
    if (nl == 1) {
        int rc = op.bottom(M->grids[0], b, x, false, 100, 1.0);
        if (rc) return rc;
        return op.bottom(M->grids[0], b, x, false, 100, 1.0);
    }
    // residuals down to the bottom (coarse iterates are zero)
    for (int l = 0; l < nl - 1; ++l) {
        int rc = op.resid_restrict(M->grids[l], (const T*)M->b[l], l == 0 ? x : nullptr, M->grids[l + 1],
                                   (T*)M->b[l + 1]);
        if (rc) return rc;
    }
    int rc = op.bottom(M->grids[nl - 1], (const T*)M->b[nl - 1], (T*)M->x[nl - 1], true, 100, 1.0);
    if (rc) return rc;
    M->cur[nl - 1] = M->x[nl - 1];
    // prolong and V-cycle every level (coarse iterates start from zero)
    for (int l = nl - 2; l >= 0; --l) {
        rc = vcycle_from<T>(M, l, l == 0 ? x : nullptr, (const T*)M->cur[l + 1], s);
        if (rc) return rc;
    }
    rc = vcycle_from<T>(M, 0, (const T*)M->cur[0], nullptr, s);
    if (rc) return rc;
    if (M->cur[0] != x) VF_HIP(hipMemcpyAsync(x, M->cur[0], M->grids[0].cells() * sizeof(T), hipMemcpyDeviceToDevice, s));
    return 0;
}

size_t elem(int type) { return type == VFLIP_MG3D_DOUBLE ? sizeof(double) : sizeof(float); }

}  // namespace

extern "C" {

int vflip_mg3d_create(int gridx, int gridy, int gridz, int type, int dirichlet, vflip_mg3d** out) {
    VF_ARG(out, "out");
    *out = nullptr;
    VF_ARG(gridx > 0 && gridy > 0 && gridz > 0, "grid extents");
    VF_ARG((size_t)gridx * gridy * gridz < ((size_t)1 << 31), "grid too large (2^31 cells)");
    VF_ARG(type == VFLIP_MG3D_FLOAT || type == VFLIP_MG3D_DOUBLE, "type");
    auto* M = new vflip_mg3d();
    M->type = type;
    M->dirichlet = dirichlet ? 1 : 0;
    M->grids = level_grids(gridx, gridy, gridz);
    const int nl = (int)M->grids.size();
    if (nl > kMaxLevels) {
        delete M;
        set_error("too many levels");
        return VFLIP_ERR_ARG;
    }
    const size_t es = elem(type);
    // x and b of levels >= 1, the ping-pong buffer of every level, 256-byte aligned
    auto pad = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    size_t total = 0;
    for (int l = 0; l < nl; ++l) total += pad(M->grids[l].cells() * es) * (l == 0 ? 1 : 3);
    hipError_t e = hipGetDevice(&M->device);
    if (e == hipSuccess) e = hipMalloc(&M->pool, total);
    if (e != hipSuccess) {
        delete M;
        set_error("vflip_mg3d_create: %s", hipGetErrorString(e));
        return VFLIP_ERR_NOMEM;
    }
    char* p = (char*)M->pool;
    M->x.assign(nl, nullptr);
    M->t.assign(nl, nullptr);
    M->b.assign(nl, nullptr);
    M->cur.assign(nl, nullptr);
    for (int l = 0; l < nl; ++l) {
        const size_t sz = pad(M->grids[l].cells() * es);
        M->t[l] = p;
        p += sz;
        if (l > 0) {
            M->x[l] = p;
            p += sz;
            M->b[l] = p;
            p += sz;
        }
    }
    *out = M;
    return 0;
}

int vflip_mg3d_destroy(vflip_mg3d* M) {
    if (!M) return 0;
    if (M->pool) (void)hipFree(M->pool);
    delete M;
    return 0;
}

int vflip_mg3d_levels(const vflip_mg3d* M, int* levels, int* dims) {
    VF_ARG(M && levels, "args");
    *levels = (int)M->grids.size();
    if (dims)
        for (size_t l = 0; l < M->grids.size(); ++l)
            for (int a = 0; a < 3; ++a) dims[3 * l + a] = M->grids[l].n[a];
    return 0;
}

int vflip_mg3d_vcycle(vflip_mg3d* M, void* x, const void* b, void* stream) {
    VF_ARG(M && x && b, "args");
    VF_ARG(x != b, "x and b must differ");
    hipStream_t s = (hipStream_t)stream;
    VF_HIP(hipSetDevice(M->device));
    int rc = M->type == VFLIP_MG3D_DOUBLE ? vcycle_top<double>(M, (double*)x, (const double*)b, s)
                                          : vcycle_top<float>(M, (float*)x, (const float*)b, s);
    if (rc) return rc;
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

int vflip_mg3d_full_multigrid(vflip_mg3d* M, void* x, const void* b, void* stream) {
    VF_ARG(M && x && b, "args");
    VF_ARG(x != b, "x and b must differ");
    hipStream_t s = (hipStream_t)stream;
    VF_HIP(hipSetDevice(M->device));
    int rc = M->type == VFLIP_MG3D_DOUBLE ? fmg_top<double>(M, (double*)x, (const double*)b, s)
                                          : fmg_top<float>(M, (float*)x, (const float*)b, s);
    if (rc) return rc;
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

int vflip_mg3d_rbgs(int nx, int ny, int nz, int type, int dirichlet, const void* b, void* x, void* x_new,
                    int iter_time, double omega, void* stream) {
    VF_ARG(nx > 0 && ny > 0 && nz > 0 && b && x && x_new && iter_time >= 0, "rbgs args");
    VF_ARG(x != x_new && x != b, "x, x_new and b must differ");
    VF_ARG(type == VFLIP_MG3D_FLOAT || type == VFLIP_MG3D_DOUBLE, "type");
    hipStream_t s = (hipStream_t)stream;
    const Grid g{{nx, ny, nz}};
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        Ops<T> op{s, dirichlet ? 1 : 0};
        if (iter_time == 0) return 0;
        if (g.cells() <= (size_t)kBottomCells) return op.bottom(g, (const T*)b, (T*)x, false, iter_time, omega);
        T* res = nullptr;
        int rc = op.smooth(g, (const T*)b, (const T*)x, nullptr, g, (T*)x, (T*)x_new, iter_time, omega, &res);
        if (rc) return rc;
        if (res != x) VF_HIP(hipMemcpyAsync(x, res, g.cells() * sizeof(T), hipMemcpyDeviceToDevice, s));
        return 0;
    };
    int rc = type == VFLIP_MG3D_DOUBLE ? run(double()) : run(float());
    if (rc) return rc;
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

int vflip_mg3d_residual(int nx, int ny, int nz, int type, int dirichlet, void* res, const void* b, const void* x,
                        void* stream) {
    VF_ARG(nx > 0 && ny > 0 && nz > 0 && res && b && x, "residual args");
    VF_ARG(type == VFLIP_MG3D_FLOAT || type == VFLIP_MG3D_DOUBLE, "type");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)nx * ny * nz;
    if (type == VFLIP_MG3D_DOUBLE)
        k_mg3d_residual<double><<<grid_for(n), kBlock, 0, s>>>(nx, ny, nz, dirichlet ? 1 : 0, (const double*)b,
                                                               (const double*)x, (double*)res);
    else
        k_mg3d_residual<float><<<grid_for(n), kBlock, 0, s>>>(nx, ny, nz, dirichlet ? 1 : 0, (const float*)b,
                                                              (const float*)x, (float*)res);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

int vflip_mg3d_restrict(int nx, int ny, int nz, int type, const void* src, void* dst, void* stream) {
    VF_ARG(nx > 0 && ny > 0 && nz > 0 && src && dst && src != dst, "restrict args");
    VF_ARG(type == VFLIP_MG3D_FLOAT || type == VFLIP_MG3D_DOUBLE, "type");
    hipStream_t s = (hipStream_t)stream;
    const Grid cg = coarser(Grid{{nx, ny, nz}});
    if (type == VFLIP_MG3D_DOUBLE)
        k_mg3d_restrict<double, false, false><<<grid_for(cg.cells()), kBlock, 0, s>>>(
            nx, ny, nz, 0, (const double*)src, nullptr, cg.n[0], cg.n[1], cg.n[2], (double*)dst);
    else
        k_mg3d_restrict<float, false, false><<<grid_for(cg.cells()), kBlock, 0, s>>>(
            nx, ny, nz, 0, (const float*)src, nullptr, cg.n[0], cg.n[1], cg.n[2], (float*)dst);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

int vflip_mg3d_prolongate(int nx, int ny, int nz, int type, const void* src, void* dst, void* stream) {
    VF_ARG(nx > 0 && ny > 0 && nz > 0 && src && dst && src != dst, "prolongate args");
    VF_ARG(type == VFLIP_MG3D_FLOAT || type == VFLIP_MG3D_DOUBLE, "type");
    hipStream_t s = (hipStream_t)stream;
    const Grid g{{nx, ny, nz}}, cg = coarser(g);
    if (type == VFLIP_MG3D_DOUBLE)
        k_mg3d_prolongate<double><<<grid_for(g.cells()), kBlock, 0, s>>>(nx, ny, nz, (const double*)src, cg.n[0],
                                                                         cg.n[1], cg.n[2], (double*)dst);
    else
        k_mg3d_prolongate<float><<<grid_for(g.cells()), kBlock, 0, s>>>(nx, ny, nz, (const float*)src, cg.n[0],
                                                                        cg.n[1], cg.n[2], (float*)dst);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
