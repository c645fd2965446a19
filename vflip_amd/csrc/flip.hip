// FLIP particle-grid transfer: cell binning, particle-to-grid, grid-to-particle
// and particle advection.
//
// Restates, for device-resident particles (xyz-interleaved float pos/vel):
//   hash_FLIP            S/Smoke_solver3D.cpp:1327-1346, S/sub_domain.cpp:1087-1107
//   particle_to_grid     S/Smoke_solver3D.cpp:1451-1510, S/sub_domain.cpp:1204-1265
//   getDelta             S/Smoke_solver3D.cpp:1297-1318
//   particle_from_grid   S/Smoke_solver3D.cpp:1398-1406 (+ PIC blend of S/fluid_particle.h:297-336)
//   advect_FLIP          S/Smoke_solver3D.cpp:141-154, 1408-1430; S/sub_domain.cpp:116-129, 1180-1202
//
// Binning is bit-exact: keys from the reference's float arithmetic, and within
// a cell particles in ascending id order (the append order of the reference's
// per-cell lists). Compiled without FMA contraction so weights, samples and
// traces round exactly as the reference's expressions do.
#pragma clang fp contract(off)

#include <climits>
#include <mutex>

#include "common.h"
#include "flip_internal.h"
#include "sampling.h"


namespace vflip {

// ----------------------------------------------------------------- binning

struct BinGeom {
    float h;
    int nx, ny, nz;
    int clamp;
    float bx, by, bz;
};

// Cell indices of a position; returns false when the coarse-grid rule drops it.
__device__ inline bool cell_of(float x, float y, float z, const BinGeom& g, int& i, int& j, int& k) {
    const float half = 0.5f * g.h;
    if (g.clamp) {
        x = x - g.bx;
        y = y - g.by;
        z = z - g.bz;
    }
    x = x + half;
    y = y + half;
    z = z + half;
    i = trunc_i32(x / g.h);
    j = trunc_i32(y / g.h);
    k = trunc_i32(z / g.h);
    if (g.clamp) {
        i = i < 0 ? 0 : (i > g.nx - 1 ? g.nx - 1 : i);
        j = j < 0 ? 0 : (j > g.ny - 1 ? g.ny - 1 : j);
        k = k < 0 ? 0 : (k > g.nz - 1 ? g.nz - 1 : k);
        return true;
    }
    return i >= 0 && j >= 0 && k >= 0 && i < g.nx && j < g.ny && k < g.nz;
}

__global__ __launch_bounds__(kBlock) void k_fill_i32(int* a, size_t n, int v) {
    for (size_t c = blockIdx.x * (size_t)kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock) a[c] = v;
}

// Key per particle and its slot among the particles of its cell. Runs of equal
// keys inside a wave (the common case for particles stored cell-major) take one
// atomic per run.
__global__ __launch_bounds__(kBlock) void k_bin_count(int np, const float* __restrict__ pos, BinGeom g,
                                                      int* __restrict__ key, int* __restrict__ slot,
                                                      int* __restrict__ count) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    int kk = -1;
    if (p < np) {
        int i, j, k;
        if (cell_of(pos[3 * (size_t)p], pos[3 * (size_t)p + 1], pos[3 * (size_t)p + 2], g, i, j, k))
            kk = (k * g.ny + j) * g.nx + i;
    }
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(kk, 1, 64);
    const bool head = lane == 0 || kk != prev;
    const unsigned long long heads = __ballot(head);
    const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int hl = 63 - __clzll((long long)(heads & le));
    const unsigned long long after = heads & ~le;
    const int nh = after ? __ffsll((unsigned long long)after) - 1 : 64;
    int base = 0;
    if (head && kk >= 0) base = atomicAdd(&count[kk], nh - lane);
    base = __shfl(base, hl, 64);
    if (p < np) {
        key[p] = kk;
        slot[p] = base + (lane - hl);
    }
}

// Exclusive scan of n ints in place, over chunks of kScanChunk.
constexpr int kScanItems = 8;
constexpr int kScanChunk = kBlock * kScanItems;

// Block-wide exclusive scan of one value per thread (kBlock threads); returns
// the exclusive prefix and the block total.
__device__ inline int block_exclusive_scan(int v, int& total) {
    __shared__ int wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) {
        if (q < wid) before += wsum[q];
        total += wsum[q];
    }
    __syncthreads();
    return before + inc - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(size_t n, const int* __restrict__ a, int* __restrict__ partial) {
    const size_t base = (size_t)blockIdx.x * kScanChunk + (size_t)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q)
        if (base + q < n) s += a[base + q];
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// One workgroup of 1024 threads: exclusive scan of the partials, grand total to *total.
__global__ __launch_bounds__(1024) void k_scan_partials(int nb, int* __restrict__ partial, int* __restrict__ total) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per;
    int s = 0;
    for (int q = 0; q < per; ++q)
        if (b0 + q < nb) s += partial[b0 + q];
    int inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int x = __shfl_up(inc, o, 64);
        if (lane >= o) inc += x;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int before = 0, all = 0;
    for (int q = 0; q < 16; ++q) {
        if (q < wid) before += wsum[q];
        all += wsum[q];
    }
    int run = before + inc - s;
    for (int q = 0; q < per; ++q)
        if (b0 + q < nb) {
            const int x = partial[b0 + q];
            partial[b0 + q] = run;
            run += x;
        }
    if (t == 0) *total = all;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(size_t n, int* __restrict__ a, const int* __restrict__ partial) {
    const size_t base = (size_t)blockIdx.x * kScanChunk + (size_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        v[q] = base + q < n ? a[base + q] : 0;
        s += v[q];
    }
    int total;
    int run = block_exclusive_scan(s, total) + partial[blockIdx.x];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q)
        if (base + q < n) {
            a[base + q] = run;
            run += v[q];
        }
}

__global__ __launch_bounds__(kBlock) void k_bin_scatter(int np, const int* __restrict__ key, const int* __restrict__ slot,
                                                        const int* __restrict__ start, int* __restrict__ order) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    const int kk = key[p];
    if (kk >= 0) order[start[kk] + slot[p]] = p;
}

// Atomic slots order a cell's particles by arrival; the reference order is
// ascending particle id. Short unsorted segments are insertion-sorted by their
// cell's thread; longer ones are queued for k_sort_long. One workgroup per
// kBlock consecutive cells: their span of `order` is staged in LDS (read
// coalesced, sorted there, written back only if a segment changed), unless it
// exceeds kFixSpan entries, when the threads work on global memory directly.
constexpr int kShortSegment = 64;
constexpr int kFixSpan = 4096;

// Sorts seg[0..n) (n <= kShortSegment) by insertion; returns whether it changed.
template <typename P> __device__ inline bool fix_segment(P seg, int n) {
    bool sorted = true;
    for (int a = 1; a < n && sorted; ++a) sorted = seg[a - 1] < seg[a];
    if (sorted) return false;
    for (int a = 1; a < n; ++a) {
        const int x = seg[a];
        int b = a - 1;
        while (b >= 0 && seg[b] > x) {
            seg[b + 1] = seg[b];
            --b;
        }
        seg[b + 1] = x;
    }
    return true;
}

__device__ inline bool segment_sorted(const int* seg, int n) {
    for (int a = 1; a < n; ++a)
        if (seg[a - 1] > seg[a]) return false;
    return true;
}

__global__ __launch_bounds__(kBlock) void k_cell_fix(size_t ncell, const int* __restrict__ start, int* __restrict__ order,
                                                     int* __restrict__ big) {
    __shared__ int s_start[kBlock + 1];
    __shared__ int s_ord[kFixSpan];
    __shared__ int s_dirty;
    const int tid = threadIdx.x;
    const size_t c0 = (size_t)blockIdx.x * kBlock;
    const int nt = (int)min((size_t)kBlock, ncell - c0);
    s_start[tid] = start[c0 + min(tid, nt)];
    if (tid == 0) {
        s_start[kBlock] = start[c0 + nt];
        s_dirty = 0;
    }
    __syncthreads();
    const int t0 = s_start[0], span = s_start[nt] - t0;
    const int s = s_start[min(tid, nt)] - t0, n = tid < nt ? s_start[tid + 1] - s_start[tid] : 0;
    if (span <= kFixSpan) {
        for (int t = tid; t < span; t += kBlock) s_ord[t] = order[t0 + t];
        __syncthreads();
        if (n >= 2) {
            if (n <= kShortSegment) {
                if (fix_segment(s_ord + s, n)) s_dirty = 1;
            } else if (!segment_sorted(s_ord + s, n)) {
                const int q = atomicAdd(&big[0], 1);
                big[1 + q] = (int)(c0 + tid);
            }
        }
        __syncthreads();
        if (s_dirty)
            for (int t = tid; t < span; t += kBlock) order[t0 + t] = s_ord[t];
        return;
    }
    if (n < 2) return;
    int* seg = order + t0 + s;
    if (n <= kShortSegment) {
        fix_segment(seg, n);
    } else if (!segment_sorted(seg, n)) {
        const int q = atomicAdd(&big[0], 1);
        big[1 + q] = (int)(c0 + tid);
    }
}

// Number of entries of the sorted run r[0..n) smaller than x.
__device__ inline int count_less(const int* r, int n, int x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (r[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Long segments: insertion-sorted runs of kShortSegment, then pairwise merges
// where every element finds its output slot by a binary search in the other
// run (ids are distinct, so ranks never tie). One workgroup per segment.
__global__ __launch_bounds__(kBlock) void k_sort_long(const int* __restrict__ start, int* __restrict__ order,
                                                      int* __restrict__ tmp, const int* __restrict__ big) {
    const int nbig = big[0];
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int c = big[1 + b];
        const int s = start[c], m = start[c + 1] - s;
        int* src = order + s;
        int* dst = tmp + s;
        for (int r0 = threadIdx.x * kShortSegment; r0 < m; r0 += kBlock * kShortSegment) {
            const int r1 = min(r0 + kShortSegment, m);
            for (int a = r0 + 1; a < r1; ++a) {
                const int x = src[a];
                int q = a - 1;
                while (q >= r0 && src[q] > x) {
                    src[q + 1] = src[q];
                    --q;
                }
                src[q + 1] = x;
            }
        }
        __syncthreads();
        for (int w = kShortSegment; w < m; w <<= 1) {
            for (int a = threadIdx.x; a < m; a += kBlock) {
                const int base = (a / (2 * w)) * (2 * w);
                const int mid = min(base + w, m), end = min(base + 2 * w, m);
                const int x = src[a];
                const int at = a < mid ? (a - base) + count_less(src + mid, end - mid, x)
                                       : (a - mid) + count_less(src + base, mid - base, x);
                dst[base + at] = x;
            }
            __syncthreads();
            int* t = src;
            src = dst;
            dst = t;
        }
        if (src != order + s) {
            for (int a = threadIdx.x; a < m; a += kBlock) order[s + a] = src[a];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------- P2G
// H(r) of S/Smoke_solver3D.h:217-223.
__device__ inline float tent(float r) {
    float res = 0;
    if (r >= -1 && r < 0) res = 1 + r;
    if (r >= 0 && r < 1) res = 1 - r;
    return res;
}

// Branch-free H(r) for the scatter paths: same values for every r.
__device__ inline float tent_nb(float r) {
    const float t = 1.f - fabsf(r);
    return (r >= -1.f && r < 1.f) ? t : 0.f;
}

// World coordinate of grid index f with offset o: ((float)f - o)*h rounded to float.
__device__ inline float grid_world(int f, double o, double hd) { return (float)(((double)(float)f - o) * hd); }

struct P2GArgs {
    const float* pos;
    const float* vel;
    const int* start;
    const int* order;
    int nx, ny, nz;
    float h;
    int sub;
    float bx, by, bz;
    float* field[3];  // scatter: sum w*v accumulators; then field values
    float* coef[3];   // scatter: sum w accumulators; then 1e-5 + sum w
};

__host__ __device__ inline int face_dim(int comp, int axis, int nx, int ny, int nz) {
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    return n + (comp == axis ? 1 : 0);
}

// Subdomain faces within 4 cells of the boundary use the wide kernel
// (SubDomainSolver::determineDist, S/sub_domain.h:89-103).
__device__ inline bool sub_shell(int i, int j, int k, int nx, int ny, int nz) {
    const int d = min(min(min(i, nx - i), min(j, ny - j)), min(k, nz - k));
    return d <= 4;
}

// Tile of cells per workgroup for the scatter; its faces plus a one-face
// halo on every side are accumulated in LDS.
constexpr int kPTX = 8, kPTY = 8, kPTZ = 8;
constexpr int kPLX = kPTX + 2, kPLY = kPTY + 2, kPLZ = kPTZ + 2, kPLN = kPLX * kPLY * kPLZ;

// LDS-binned scatter: each workgroup walks the particles of its tile's cells
// (rows of cells are contiguous in the binned order), accumulates the tent
// weights of all three components into LDS with LDS atomics, then adds its
// non-zero LDS faces to the global accumulators with one global atomic each.
__global__ __launch_bounds__(kBlock) void k_p2g_scatter(P2GArgs a, int tiles_x, int tiles_y, int ntiles) {
    __shared__ float acc[3][2][kPLN];
    for (int q = threadIdx.x; q < 3 * 2 * kPLN; q += kBlock) (&acc[0][0][0])[q] = 0.f;
    __syncthreads();
    // XCD-aware tile order: consecutive workgroups go round-robin over the 8
    // XCDs, so hand each XCD a contiguous band of tiles (shared halo faces stay
    // in one L2).
    const int L = blockIdx.x;
    int t = L;
    if (ntiles % 8 == 0) t = (L % 8) * (ntiles / 8) + L / 8;
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tz = t / (tiles_x * tiles_y);
    const int x0 = tx * kPTX, y0 = ty * kPTY, z0 = tz * kPTZ;
    const int x1 = min(x0 + kPTX, a.nx);
    const double hd = a.h;
    const BinGeom g{a.h, a.nx, a.ny, a.nz, a.sub, a.bx, a.by, a.bz};
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int r = wid; r < kPTY * kPTZ; r += kBlock / 64) {
        const int y = y0 + r % kPTY, z = z0 + r / kPTY;
        if (y >= a.ny || z >= a.nz) continue;
        const size_t c0 = lin3(x0, y, z, a.nx, a.ny);
        const int q0 = a.start[c0], q1 = a.start[c0 + (x1 - x0)];
        for (int q = q0 + lane; q < q1; q += 64) {
            const int p = a.order[q];
            const float px = a.pos[3 * (size_t)p], py = a.pos[3 * (size_t)p + 1], pz = a.pos[3 * (size_t)p + 2];
            const float vv[3] = {a.vel[3 * (size_t)p], a.vel[3 * (size_t)p + 1], a.vel[3 * (size_t)p + 2]};
            int ci[3];
            cell_of(px, py, pz, g, ci[0], ci[1], ci[2]);
            // the binning placed p in this tile; a caller whose bmin differs from
            // the one it binned with gets the particle skipped, not an LDS overrun
            if ((unsigned)(ci[0] - x0) >= (unsigned)kPTX || (unsigned)(ci[1] - y0) >= (unsigned)kPTY ||
                (unsigned)(ci[2] - z0) >= (unsigned)kPTZ)
                continue;
            float pp[3] = {px, py, pz};
            if (a.sub) {
                pp[0] = px - a.bx;
                pp[1] = py - a.by;
                pp[2] = pz - a.bz;
            }
            // per axis: weights of the three candidate faces ci-1..ci+1, for
            // cell-centred (offset 0) and staggered (offset .5) positions
            float wc[3][3], ws[3][3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int f = ci[ax] - 1 + d;
                    wc[ax][d] = tent((pp[ax] - grid_world(f, 0.0, hd)) / a.h);
                    ws[ax][d] = tent((pp[ax] - grid_world(f, 0.5, hd)) / a.h);
                }
            const int lx = ci[0] - 1 - (x0 - 1), ly = ci[1] - 1 - (y0 - 1), lz = ci[2] - 1 - (z0 - 1);
#pragma unroll
            for (int comp = 0; comp < 3; ++comp) {
                if (!a.field[comp]) continue;
#pragma unroll
                for (int dz = 0; dz < 3; ++dz) {
                    const float wz = comp == 2 ? ws[2][dz] : wc[2][dz];
#pragma unroll
                    for (int dy = 0; dy < 3; ++dy) {
                        const float wy = comp == 1 ? ws[1][dy] : wc[1][dy];
#pragma unroll
                        for (int dx = 0; dx < 3; ++dx) {
                            const float wx = comp == 0 ? ws[0][dx] : wc[0][dx];
                            const float wgt = wx * wy * wz;
                            if (wgt != 0.f) {
                                const int li = (lx + dx) + kPLX * ((ly + dy) + kPLY * (lz + dz));
                                atomicAdd(&acc[comp][0][li], wgt);
                                atomicAdd(&acc[comp][1][li], wgt * vv[comp]);
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int comp = 0; comp < 3; ++comp) {
        if (!a.field[comp]) continue;
        const int fx = face_dim(comp, 0, a.nx, a.ny, a.nz), fy = face_dim(comp, 1, a.nx, a.ny, a.nz),
                  fz = face_dim(comp, 2, a.nx, a.ny, a.nz);
        for (int li = threadIdx.x; li < kPLN; li += kBlock) {
            const float sw = acc[comp][0][li];
            if (sw == 0.f) continue;
            const int i = x0 - 1 + li % kPLX, j = y0 - 1 + (li / kPLX) % kPLY, k = z0 - 1 + li / (kPLX * kPLY);
            if (i < 0 || j < 0 || k < 0 || i >= fx || j >= fy || k >= fz) continue;
            if (a.sub && sub_shell(i, j, k, a.nx, a.ny, a.nz)) continue;
            const size_t o = lin3(i, j, k, fx, fy);
            unsafeAtomicAdd(&a.coef[comp][o], sw);
            unsafeAtomicAdd(&a.field[comp][o], acc[comp][1][li]);
        }
    }
}

// coef = 1e-5 + sum w, field = sum w*v / coef (S/Smoke_solver3D.cpp:1475-1503).
__global__ __launch_bounds__(kBlock) void k_p2g_normalize(int comp, int nx, int ny, int nz, int sub,
                                                          float* __restrict__ field, float* __restrict__ coef) {
    const int fx = face_dim(comp, 0, nx, ny, nz), fy = face_dim(comp, 1, nx, ny, nz), fz = face_dim(comp, 2, nx, ny, nz);
    const size_t n = (size_t)fx * fy * fz;
    for (size_t o = blockIdx.x * (size_t)kBlock + threadIdx.x; o < n; o += (size_t)gridDim.x * kBlock) {
        if (sub) {
            const int i = (int)(o % fx), j = (int)((o / fx) % fy), k = (int)(o / ((size_t)fx * fy));
            if (sub_shell(i, j, k, nx, ny, nz)) continue;
        }
        const float tw = 1e-5f + coef[o];
        coef[o] = tw;
        field[o] = field[o] / tw;
    }
}

// Owner-computes P2G. A workgroup owns the faces of a kOT^3 block of face
// indices (the last block along an axis also owns the extra staggered face).
// Every cell within one cell of the block is walked by one thread, which sums
// its particles' weights in registers, one component at a time: the staggered
// axis reaches faces c and c+1 (c-1 only when rounding puts a particle a hair
// below its cell's lower face, handled apart with LDS atomics) and the other
// axes faces c-1..c+1, so 2 x 3 x 3 faces. The sums go to LDS (one add per face
// per cell; the lanes of one add target distinct faces), and the block then
// writes coef = 1e-5 + sum w and field = sum w*v / coef with plain stores: no
// global atomics, no zero fill, no separate normalisation pass. One component
// per pass keeps the register sums at 36, so enough waves stay resident to hide
// the particle loads and the LDS adds.
constexpr int kOT = 8;
constexpr int kOH = kOT + 2;  // cells per axis that touch the block's faces
constexpr int kOF = kOT + 1;  // LDS faces per axis (extra face of the last block)
constexpr int kOFN = kOF * kOF * kOF;

struct OwnerTile {
    int b0[3];
    int hi[3][3];  // owned face range [b0, hi) per component and axis
};

// Particles of one cell held in registers: the first kOwnChunk are loaded
// together (independent loads in flight at once) and reused by all three
// components; cells holding more are rare and read the rest per component.
constexpr int kOwnChunk = 8;

struct CellParticles {
    float r[kOwnChunk][3];  // scaled offsets to the centred face c-1 per axis
    float v[kOwnChunk][3];
};

// One scaled offset per axis: r = distance in cells from the centred face c-1;
// the other candidate faces sit whole or half cells from it.
__device__ inline void scaled_offsets(const P2GArgs& a, size_t p, const float (&g0)[3], float inv_h, float (&r)[3]) {
    const float bsh[3] = {a.bx, a.by, a.bz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float x = a.pos[3 * p + ax];
        if (a.sub) x = x - bsh[ax];
        r[ax] = (x - g0[ax]) * inv_h;
    }
}

// Rare: a particle a rounding hair below its cell's lower staggered face puts
// face c-1 inside the tent support. Kept out of line so the hot loop stays small.
template <int C>
__device__ __attribute__((noinline)) void owner_wedge(int b0x, int b0y, int b0z, int hix, int hiy, int hiz, int cx,
                                                      int cy, int cz, float wedge, float wu0, float wu1, float wu2,
                                                      float wv0, float wv1, float wv2, float vc,
                                                      float2 (&acc)[3][kOFN]) {
    constexpr int UA = C == 0 ? 1 : 0, VA = C == 2 ? 1 : 2;
    const int ci[3] = {cx, cy, cz}, b0[3] = {b0x, b0y, b0z}, hi[3] = {hix, hiy, hiz};
    const float wu[3] = {wu0, wu1, wu2}, wv[3] = {wv0, wv1, wv2};
    for (int u = 0; u < 3; ++u)
        for (int v = 0; v < 3; ++v) {
            int fi[3];
            fi[C] = ci[C] - 1;
            fi[UA] = ci[UA] - 1 + u;
            fi[VA] = ci[VA] - 1 + v;
            const float fx = C == 0 ? wedge : wu[u];
            const float fy = C == 1 ? wedge : (UA == 1 ? wu[u] : wv[v]);
            const float fz = C == 2 ? wedge : wv[v];
            const float wgt = fx * fy * fz;
            bool own = wgt != 0.f;
            for (int ax = 0; ax < 3; ++ax) own = own && fi[ax] >= b0[ax] && fi[ax] < hi[ax];
            if (own) {
                const int li = (fi[0] - b0[0]) + kOF * ((fi[1] - b0[1]) + kOF * (fi[2] - b0[2]));
                atomicAdd(&acc[C][li].x, wgt);
                atomicAdd(&acc[C][li].y, wgt * vc);
            }
        }
}

template <int C>
__device__ inline void owner_accumulate(const P2GArgs& a, const OwnerTile& T, const int (&ci)[3], const float (&r)[3],
                                        float vc, float (&sw)[2][3][3], float (&sv)[2][3][3],
                                        float2 (&acc)[3][kOFN]) {
    constexpr int UA = C == 0 ? 1 : 0, VA = C == 2 ? 1 : 2;  // the two centred axes, UA < VA
    const float ws[2] = {tent_nb(r[C] - 0.5f), tent_nb(r[C] - 1.5f)};
    const float wu[3] = {tent_nb(r[UA]), tent_nb(r[UA] - 1.f), tent_nb(r[UA] - 2.f)};
    const float wv[3] = {tent_nb(r[VA]), tent_nb(r[VA] - 1.f), tent_nb(r[VA] - 2.f)};
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                // x*y*z factor order of compute_weight
                const float fx = C == 0 ? ws[s] : wu[u];
                const float fy = C == 1 ? ws[s] : (UA == 1 ? wu[u] : wv[v]);
                const float fz = C == 2 ? ws[s] : wv[v];
                const float wgt = fx * fy * fz;
                sw[s][u][v] += wgt;
                sv[s][u][v] = __builtin_fmaf(wgt, vc, sv[s][u][v]);
            }
    const float wedge = tent_nb(r[C] + 0.5f);
    if (wedge != 0.f)
        owner_wedge<C>(T.b0[0], T.b0[1], T.b0[2], T.hi[C][0], T.hi[C][1], T.hi[C][2], ci[0], ci[1], ci[2], wedge, wu[0],
                       wu[1], wu[2], wv[0], wv[1], wv[2], vc, acc);
}

// Component C of one cell: the register-held particles, then any beyond
// kOwnChunk. Cells past n (or lanes without a cell) hold offsets outside every
// tent and add exact zeros.
template <int C>
__device__ inline void owner_sums(const P2GArgs& a, const OwnerTile& T, const int (&ci)[3], int q0, int n,
                                  const CellParticles& cp, const float (&g0)[3], float inv_h, float (&sw)[2][3][3],
                                  float (&sv)[2][3][3], float2 (&acc)[3][kOFN]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
            for (int v = 0; v < 3; ++v) sw[s][u][v] = sv[s][u][v] = 0.f;
#pragma unroll
    for (int k = 0; k < kOwnChunk; ++k) owner_accumulate<C>(a, T, ci, cp.r[k], cp.v[k][C], sw, sv, acc);
    for (int k = kOwnChunk; k < n; ++k) {
        const size_t p = (size_t)a.order[q0 + k];
        float r[3];
        scaled_offsets(a, p, g0, inv_h, r);
        owner_accumulate<C>(a, T, ci, r, a.vel[3 * p + C], sw, sv, acc);
    }
}

// Adds the cell sums to the tile's faces with plain LDS read-modify-writes in
// 18 barrier-separated phases, one per relative face offset: within a phase
// every thread's face is distinct (one cell per thread), so no atomics are
// needed. Every thread of the workgroup takes part (the barriers are uniform).
template <int C>
__device__ inline void owner_flush(const OwnerTile& T, const int (&ci)[3], bool valid, const float (&sw)[2][3][3],
                                   const float (&sv)[2][3][3], float2 (&acc)[3][kOFN]) {
    constexpr int UA = C == 0 ? 1 : 0, VA = C == 2 ? 1 : 2;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                int fi[3];
                fi[C] = ci[C] + s;
                fi[UA] = ci[UA] - 1 + u;
                fi[VA] = ci[VA] - 1 + v;
                bool own = valid;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) own = own && fi[ax] >= T.b0[ax] && fi[ax] < T.hi[C][ax];
                if (own) {
                    const int li = (fi[0] - T.b0[0]) + kOF * ((fi[1] - T.b0[1]) + kOF * (fi[2] - T.b0[2]));
                    float2 x = acc[C][li];
                    x.x += sw[s][u][v];
                    x.y += sv[s][u][v];
                    acc[C][li] = x;
                }
                __syncthreads();
            }
}

__global__ __launch_bounds__(kBlock) void k_p2g_owner(P2GArgs a, int tiles_x, int tiles_y, int ntiles) {
    __shared__ float2 acc[3][kOFN];
    const int tid = threadIdx.x;
    for (int q = tid; q < 3 * kOFN; q += kBlock) (&acc[0][0])[q] = make_float2(0.f, 0.f);
    const int L = blockIdx.x;
    int t = L;
    if (ntiles % 8 == 0) t = (L % 8) * (ntiles / 8) + L / 8;
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tz = t / (tiles_x * tiles_y);
    OwnerTile T;
    T.b0[0] = tx * kOT;
    T.b0[1] = ty * kOT;
    T.b0[2] = tz * kOT;
    const int nn[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int fd = nn[ax] + (c == ax ? 1 : 0);
            T.hi[c][ax] = T.b0[ax] + kOT >= nn[ax] ? fd : T.b0[ax] + kOT;
        }
    const double hd = a.h;
    const float inv_h = 1.0f / a.h;
    for (int base = 0; base < kOH * kOH * kOH; base += kBlock) {
        const int cid = base + tid;
        const int ci[3] = {T.b0[0] - 1 + cid % kOH, T.b0[1] - 1 + (cid / kOH) % kOH, T.b0[2] - 1 + cid / (kOH * kOH)};
        int q0 = 0, n = 0;
        if (cid < kOH * kOH * kOH && ci[0] >= 0 && ci[1] >= 0 && ci[2] >= 0 && ci[0] < a.nx && ci[1] < a.ny &&
            ci[2] < a.nz) {
            const size_t cell = lin3(ci[0], ci[1], ci[2], a.nx, a.ny);
            q0 = a.start[cell];
            n = a.start[cell + 1] - q0;
        }
        // uniform across the workgroup: skip rounds without particles
        if (!__syncthreads_or(n > 0)) continue;
        const bool valid = n > 0;
        const float g0[3] = {grid_world(ci[0] - 1, 0.0, hd), grid_world(ci[1] - 1, 0.0, hd),
                             grid_world(ci[2] - 1, 0.0, hd)};
        CellParticles cp;
        int pid[kOwnChunk];
#pragma unroll
        for (int k = 0; k < kOwnChunk; ++k) pid[k] = k < n ? a.order[q0 + k] : -1;
#pragma unroll
        for (int k = 0; k < kOwnChunk; ++k) {
            if (pid[k] >= 0) {
                scaled_offsets(a, (size_t)pid[k], g0, inv_h, cp.r[k]);
#pragma unroll
                for (int c = 0; c < 3; ++c) cp.v[k][c] = a.vel[3 * (size_t)pid[k] + c];
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    cp.r[k][c] = 8.f;
                    cp.v[k][c] = 0.f;
                }
            }
        }
        float sw[2][3][3], sv[2][3][3];
        if (a.field[0]) {
            owner_sums<0>(a, T, ci, q0, n, cp, g0, inv_h, sw, sv, acc);
            owner_flush<0>(T, ci, valid, sw, sv, acc);
        }
        if (a.field[1]) {
            owner_sums<1>(a, T, ci, q0, n, cp, g0, inv_h, sw, sv, acc);
            owner_flush<1>(T, ci, valid, sw, sv, acc);
        }
        if (a.field[2]) {
            owner_sums<2>(a, T, ci, q0, n, cp, g0, inv_h, sw, sv, acc);
            owner_flush<2>(T, ci, valid, sw, sv, acc);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (!a.field[c]) continue;
        const int fx = face_dim(c, 0, a.nx, a.ny, a.nz), fy = face_dim(c, 1, a.nx, a.ny, a.nz);
        const int ex = T.hi[c][0] - T.b0[0], ey = T.hi[c][1] - T.b0[1], ez = T.hi[c][2] - T.b0[2];
        for (int li = tid; li < ex * ey * ez; li += kBlock) {
            const int lx = li % ex, ly = (li / ex) % ey, lz = li / (ex * ey);
            const int i = T.b0[0] + lx, j = T.b0[1] + ly, k = T.b0[2] + lz;
            if (a.sub && sub_shell(i, j, k, a.nx, a.ny, a.nz)) continue;
            const float2 x = acc[c][lx + kOF * (ly + kOF * lz)];
            const float tw = 1e-5f + x.x;
            const size_t o = lin3(i, j, k, fx, fy);
            a.coef[c][o] = tw;
            a.field[c][o] = x.y / tw;
        }
    }
}

// Per-face gather in the reference's order: cells of the (2l+1)^3 window
// z-outer/x-inner, particles in binned order, float accumulation from 1e-5.
// shell_only: only the subdomain's wide-kernel faces (the scatter did the rest).
__global__ __launch_bounds__(kBlock) void k_p2g_gather(P2GArgs a, int comp, int shell_only) {
    const int fx = face_dim(comp, 0, a.nx, a.ny, a.nz), fy = face_dim(comp, 1, a.nx, a.ny, a.nz),
              fz = face_dim(comp, 2, a.nx, a.ny, a.nz);
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= fx || j >= fy || k >= fz) return;
    int l = 1;
    float hk = a.h;
    if (a.sub && sub_shell(i, j, k, a.nx, a.ny, a.nz)) {
        l = 4;
        hk = 4.0f * a.h;
    } else if (shell_only) {
        return;
    }
    const double hd = a.h;
    const float gx = grid_world(i, comp == 0 ? 0.5 : 0.0, hd), gy = grid_world(j, comp == 1 ? 0.5 : 0.0, hd),
                gz = grid_world(k, comp == 2 ? 0.5 : 0.0, hd);
    float tw = 1e-5f, tv = 0.f;
    for (int c = k - l; c <= k + l; ++c) {
        if (c < 0 || c >= a.nz) continue;
        for (int b = j - l; b <= j + l; ++b) {
            if (b < 0 || b >= a.ny) continue;
            for (int ia = i - l; ia <= i + l; ++ia) {
                if (ia < 0 || ia >= a.nx) continue;
                const size_t cell = lin3(ia, b, c, a.nx, a.ny);
                const int q1 = a.start[cell + 1];
                for (int q = a.start[cell]; q < q1; ++q) {
                    const int p = a.order[q];
                    float px = a.pos[3 * (size_t)p], py = a.pos[3 * (size_t)p + 1], pz = a.pos[3 * (size_t)p + 2];
                    if (a.sub) {
                        px = px - a.bx;
                        py = py - a.by;
                        pz = pz - a.bz;
                    }
                    const float wgt = tent((px - gx) / hk) * tent((py - gy) / hk) * tent((pz - gz) / hk);
                    tw += wgt;
                    tv += wgt * a.vel[3 * (size_t)p + comp];
                }
            }
        }
    }
    const size_t o = lin3(i, j, k, fx, fy);
    a.coef[comp][o] = tw;
    a.field[comp][o] = tv / tw;
}

// ---------------------------------------------------------------------- G2P
__global__ __launch_bounds__(kBlock) void k_get_delta(size_t n, const float* __restrict__ fnew,
                                                      const float* __restrict__ fold, float* __restrict__ out) {
    for (size_t c = blockIdx.x * (size_t)kBlock + threadIdx.x; c < n; c += (size_t)gridDim.x * kBlock)
        out[c] = fnew[c] - fold[c];
}

struct G2PArgs {
    int nx, ny, nz;
    float h;
    const float* fnew[3];
    const float* fold[3];  // all NULL: fnew holds the delta grids
    float flip;
    float bx, by, bz;
};

__global__ __launch_bounds__(kBlock) void k_g2p(int np, const float* __restrict__ pos, float* __restrict__ vel,
                                                G2PArgs a) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    const double hd = a.h;
    const float x = pos[3 * (size_t)p] - a.bx, y = pos[3 * (size_t)p + 1] - a.by, z = pos[3 * (size_t)p + 2] - a.bz;
    const int dims[3][3] = {{a.nx + 1, a.ny, a.nz}, {a.nx, a.ny + 1, a.nz}, {a.nx, a.ny, a.nz + 1}};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double ox = c == 0 ? 0.5 : 0.0, oy = c == 1 ? 0.5 : 0.0, oz = c == 2 ? 0.5 : 0.0;
        float d;
        if (a.fold[c]) {
            const DeltaGrid dg{a.fnew[c], a.fold[c], dims[c][0], dims[c][1], dims[c][2], ox, oy, oz};
            d = sample_linear(dg, hd, x, y, z);
        } else {
            const Grid dg{a.fnew[c], dims[c][0], dims[c][1], dims[c][2], ox, oy, oz};
            d = sample_linear(dg, hd, x, y, z);
        }
        float& vc = vel[3 * (size_t)p + c];
        if (a.flip >= 1.0f) {
            vc += d;
        } else {
            const Grid gn{a.fnew[c], dims[c][0], dims[c][1], dims[c][2], ox, oy, oz};
            const float pic = sample_linear(gn, hd, x, y, z);
            const double fw = (double)a.flip, pw = 1.0 - fw;
            vc = (float)(pw * pic + fw * (vc + d));
        }
    }
}

// ------------------------------------------------------- particle advection
struct AdvArgs {
    int nx, ny, nz;
    float h, dt;
    int sub;
    float bx, by, bz;
    const float* u;
    const float* v;
    const float* w;
    const float* phi;
};

__device__ inline void velocity_at(const AdvArgs& a, const float* x, float* o) {
    float px = x[0], py = x[1], pz = x[2];
    if (a.sub) {
        px = px - a.bx;
        py = py - a.by;
        pz = pz - a.bz;
    }
    const double hd = a.h;
    const Grid gu{a.u, a.nx + 1, a.ny, a.nz, 0.5, 0.0, 0.0}, gv{a.v, a.nx, a.ny + 1, a.nz, 0.0, 0.5, 0.0},
        gw{a.w, a.nx, a.ny, a.nz + 1, 0.0, 0.0, 0.5};
    o[0] = sample_linear(gu, hd, px, py, pz);
    o[1] = sample_linear(gv, hd, px, py, pz);
    o[2] = sample_linear(gw, hd, px, py, pz);
}

__global__ __launch_bounds__(kBlock) void k_advect_particles(int np, float* __restrict__ pos, AdvArgs a) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= np) return;
    const float dt = a.dt;
    const float c1 = (float)(2.0 / 9.0 * dt), c2 = (float)(3.0 / 9.0 * dt);
    const float c3 = a.sub ? (float)(4.0 / 9.0 * dt) : (float)((1.0 - c1 - c2) * dt);
    const float t1 = (float)(0.5 * dt), t2 = (float)(0.75 * dt);
    float x[3] = {pos[3 * (size_t)p], pos[3 * (size_t)p + 1], pos[3 * (size_t)p + 2]};
    float v1[3], v2[3], v3[3], m[3];
    velocity_at(a, x, v1);
#pragma unroll
    for (int q = 0; q < 3; ++q) m[q] = x[q] + t1 * v1[q];
    velocity_at(a, m, v2);
#pragma unroll
    for (int q = 0; q < 3; ++q) m[q] = x[q] + t2 * v2[q];
    velocity_at(a, m, v3);
#pragma unroll
    for (int q = 0; q < 3; ++q) x[q] = x[q] + c1 * v1[q] + c2 * v2[q] + c3 * v3[q];
    const double hd = a.h;
    const Grid gp{a.phi, a.nx + 1, a.ny + 1, a.nz + 1, 0.5, 0.5, 0.5};
    const float ph = sample_linear(gp, hd, x[0], x[1], x[2]);
    if (ph < 0) {
        float n[3];
        sample_grad(gp, hd, x[0], x[1], x[2], n);
        // normalize: n /= (float)(sqrt((double)|n|^2) + 1e-6)  (S/vec.h:194-203,218-219)
        float l = n[0] * n[0];
        l += n[1] * n[1];
        l += n[2] * n[2];
        const float mg = (float)(sqrt((double)l) + 1e-6);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            n[q] /= mg;
            x[q] -= ph * n[q];
        }
    }
    pos[3 * (size_t)p] = x[0];
    pos[3 * (size_t)p + 1] = x[1];
    pos[3 * (size_t)p + 2] = x[2];
}

// ------------------------------------------------------------ reorder_FLIP
// Cell-major rebuild of the particle set (S/Smoke_solver3D.cpp:1347-1396,
// S/sub_domain.cpp:1110-1159). A cell's candidates are its first `keep`
// binned particles (keep = n < 10 ? n : 8) followed by max(8 - n, 0) new
// particles; a candidate survives when the nodal solid phi at it is >= 0.
// k_reorder_count stores each cell's survivor count and survivor bit mask
// (at most 9 candidates, bit q = candidate q), a scan turns the counts into
// output offsets, and k_reorder_emit writes the survivors: a survivor's output
// slot is its cell's offset plus the survivors before it in the mask, which is
// the reference's order (kept particles, then new ones, cell after cell).
//
// One workgroup per kReoTile consecutive cells, a lane per (cell, candidate)
// pair: consecutive lanes read consecutive entries of `order` and, once the
// particles are cell-sorted (after the first rebuild), consecutive particles,
// and write consecutive output slots. Cells whose every reachable solid-phi node
// is >= 0 keep all candidates without sampling.
struct ReoArgs {
    int nx, ny, nz;
    float h;
    int sub;  // SubDomainSolver form: interior cells only, bmin-relative sampling
    float bx, by, bz;
    unsigned seed;
    const float* pos;
    const float* vel;
    const int* start;
    const int* order;
    const float* u;
    const float* v;
    const float* w;
    const float* phi;
};


// New particle `slot` of cell c = (i, j, k): (bmin +) (i,j,k) h + (r * 0.49) h,
// key 3 (8 c + slot) + axis.
__device__ inline void reseed_pos(const ReoArgs& a, size_t c, int i, int j, int k, int slot, float* x) {
    const int ijk[3] = {i, j, k};
    const float b[3] = {a.bx, a.by, a.bz};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float r = reseed_jitter(a.seed, (unsigned)(3 * (8 * c + (size_t)slot) + ax));
        const float cell = (float)ijk[ax] * a.h;
        const float base = a.sub ? b[ax] + cell : cell;
        x[ax] = base + (r * 0.49f) * a.h;
    }
}

// get_solid_phi(pos) >= 0: nodal SDF (nx+1, ny+1, nz+1), offsets 0.5.
__device__ inline bool reo_solid_ok(const ReoArgs& a, float x, float y, float z) {
    if (a.sub) {
        x = x - a.bx;
        y = y - a.by;
        z = z - a.bz;
    }
    const Grid gp{a.phi, a.nx + 1, a.ny + 1, a.nz + 1, 0.5, 0.5, 0.5};
    return sample_linear(gp, (double)a.h, x, y, z) >= 0.0f;
}

// Solid-free cells. Every node that the solid-phi sample of a particle in cell
// (i,j,k) can read being >= 0 makes every such sample >= 0 (trilinear weights
// lie in [0, 1]), so no particle or new particle of the cell is dropped and
// none needs sampling. A particle binned in cell i has x/h + 1/2 within
// rounding of [i, i + 1), so the sampler's base node is i - 1, i or i + 1:
// nodes i-1..i+2 per axis, clamped as the sampler clamps. NaN nodes fail.
// reo_col_free tests the 4 x 4 (y, z) nodes of one x-node column; a cell is
// free when the columns at x-nodes i-1..i+2 are.
__device__ inline bool reo_col_free(const ReoArgs& a, int ii, int j, int k) {
    ii = min(max(ii, 0), a.nx);
    bool free = true;
#pragma unroll
    for (int dz = -1; dz <= 2; ++dz) {
        const int kk = min(max(k + dz, 0), a.nz);
#pragma unroll
        for (int dy = -1; dy <= 2; ++dy) {
            const int jj = min(max(j + dy, 0), a.ny);
            free &= a.phi[((size_t)kk * (a.ny + 1) + jj) * (a.nx + 1) + ii] >= 0.0f;
        }
    }
    return free;
}

// Cell coordinates and whether the cell takes part (all cells; the subdomain
// form only the interior 1..n-2).
__device__ inline bool reo_cell(const ReoArgs& a, size_t c, int& i, int& j, int& k) {
    const unsigned plane = (unsigned)a.nx * (unsigned)a.ny, cu = (unsigned)c;  // cell counts fit in int
    k = (int)(cu / plane);
    const int r = (int)(cu - (unsigned)k * plane);
    j = r / a.nx;
    i = r - j * a.nx;
    if (!a.sub) return true;
    return i >= 1 && i < a.nx - 1 && j >= 1 && j < a.ny - 1 && k >= 1 && k < a.nz - 1;
}

// Cells per workgroup (a multiple of kBlock).
constexpr int kReoTile = kBlock;
constexpr int kReoPer = kReoTile / kBlock;
// Candidates per cell: n < 8 gives n kept + (8 - n) new, 8 or 9 gives n, more
// gives 8, so at most 9. Candidate s of a cell is its binned entry s when
// s < keep, else new particle s - keep; it is also the cell's mask bit.
constexpr int kReoSlots = 9;

// Tile of cells [c0, c0 + nt): cell starts in s_start[0..kReoTile].
__device__ inline int reo_load_starts(const ReoArgs& a, size_t ncell, int* s_start) {
    const size_t c0 = (size_t)blockIdx.x * kReoTile;
    const int nt = (int)min((size_t)kReoTile, ncell - c0);
    for (int u = threadIdx.x; u <= kReoTile; u += kBlock) s_start[u] = a.start[c0 + min(u, nt)];
    return nt;
}

__device__ inline int reo_keep(int n) { return n < 10 ? n : 8; }
__device__ inline int reo_cand(int n) { return n < 8 ? 8 : reo_keep(n); }

__global__ __launch_bounds__(kBlock) void k_reorder_count(size_t ncell, ReoArgs a, int* __restrict__ cnt,
                                                          int* __restrict__ mask) {
    __shared__ int s_start[kReoTile + 1];
    __shared__ int s_mask[kReoTile];
    __shared__ int s_ijk[3][kReoTile];
    __shared__ unsigned char s_col[kReoTile];
    __shared__ unsigned char s_test[kReoTile];  // cell takes part and needs the solid test
    const int tid = threadIdx.x;
    const size_t c0 = (size_t)blockIdx.x * kReoTile;
    const int nt = reo_load_starts(a, ncell, s_start);
    // Own x-node column of every cell; the columns at i-1, i+1, i+2 come from
    // the neighbouring cells of the same row in this tile, else are read here
    // directly (node -1 clamps to node 0, the cell's own column).
    int ci[kReoPer], cj[kReoPer], ck[kReoPer];
    bool part[kReoPer], edge[kReoPer];
#pragma unroll
    for (int u = 0; u < kReoPer; ++u) {
        const int l = tid + u * kBlock;
        s_mask[l] = 0;
        ci[u] = cj[u] = ck[u] = 0;
        part[u] = l < nt && reo_cell(a, c0 + l, ci[u], cj[u], ck[u]);
        s_ijk[0][l] = ci[u];
        s_ijk[1][l] = cj[u];
        s_ijk[2][l] = ck[u];
        s_col[l] = l < nt && reo_col_free(a, ci[u], cj[u], ck[u]);
        bool e = true;
        if (part[u]) {
            if (ci[u] > 0 && l == 0) e = e && reo_col_free(a, ci[u] - 1, cj[u], ck[u]);
            if (!(ci[u] + 1 < a.nx && l + 1 < nt)) e = e && reo_col_free(a, ci[u] + 1, cj[u], ck[u]);
            if (!(ci[u] + 2 < a.nx && l + 2 < nt)) e = e && reo_col_free(a, ci[u] + 2, cj[u], ck[u]);
        }
        edge[u] = e;
    }
    __syncthreads();
    bool free[kReoPer];
#pragma unroll
    for (int u = 0; u < kReoPer; ++u) {
        const int l = tid + u * kBlock;
        bool f = part[u] && s_col[l] && edge[u];
        if (f && ci[u] > 0 && l > 0) f = s_col[l - 1];
        if (f && ci[u] + 1 < a.nx && l + 1 < nt) f = s_col[l + 1];
        if (f && ci[u] + 2 < a.nx && l + 2 < nt) f = s_col[l + 2];
        free[u] = f;
        s_test[l] = part[u] && !f;
    }
    __syncthreads();
    // a lane per (cell, candidate) pair of the cells that need the test
    for (int r = tid; r < nt * kReoSlots; r += kBlock) {
        const int l = r / kReoSlots, sl = r - l * kReoSlots;
        const int n = s_start[l + 1] - s_start[l], keep = reo_keep(n);
        if (sl >= reo_cand(n) || !s_test[l]) continue;
        float x[3];
        if (sl < keep) {
            const size_t p = (size_t)a.order[s_start[l] + sl];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) x[ax] = a.pos[3 * p + ax];
        } else {
            reseed_pos(a, c0 + l, s_ijk[0][l], s_ijk[1][l], s_ijk[2][l], sl - keep, x);
        }
        if (reo_solid_ok(a, x[0], x[1], x[2])) atomicOr(&s_mask[l], 1 << sl);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kReoPer; ++u) {
        const int l = tid + u * kBlock;
        if (l >= nt) continue;
        const int n = s_start[l + 1] - s_start[l];
        const int m = !part[u] ? 0 : (free[u] ? (1 << reo_cand(n)) - 1 : s_mask[l]);
        cnt[c0 + l] = __popc(m);
        mask[c0 + l] = m;
    }
}

__global__ __launch_bounds__(kBlock) void k_reorder_emit(size_t ncell, ReoArgs a, const int* __restrict__ offset,
                                                         const int* __restrict__ mask, float* __restrict__ out_pos,
                                                         float* __restrict__ out_vel) {
    __shared__ int s_start[kReoTile + 1];
    __shared__ int s_mask[kReoTile];
    __shared__ int s_off[kReoTile];
    const int tid = threadIdx.x;
    const size_t c0 = (size_t)blockIdx.x * kReoTile;
    const int nt = reo_load_starts(a, ncell, s_start);
    for (int l = tid; l < kReoTile; l += kBlock) {
        s_mask[l] = l < nt ? mask[c0 + l] : 0;
        s_off[l] = l < nt ? offset[c0 + l] : 0;
    }
    __syncthreads();
    const double hd = a.h;
    const Grid gu{a.u, a.nx + 1, a.ny, a.nz, 0.5, 0.0, 0.0};
    const Grid gv{a.v, a.nx, a.ny + 1, a.nz, 0.0, 0.5, 0.0};
    const Grid gw{a.w, a.nx, a.ny, a.nz + 1, 0.0, 0.0, 0.5};
    // a lane per (cell, candidate) pair; survivors go to the cell's offset plus
    // the survivors before them
    for (int r = tid; r < nt * kReoSlots; r += kBlock) {
        const int l = r / kReoSlots, sl = r - l * kReoSlots;
        const int m = s_mask[l];
        if (!((m >> sl) & 1)) continue;
        const int n = s_start[l + 1] - s_start[l], keep = reo_keep(n);
        const size_t o = (size_t)s_off[l] + __popc(m & ((1 << sl) - 1));
        if (sl < keep) {
            const size_t p = (size_t)a.order[s_start[l] + sl];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                out_pos[3 * o + ax] = a.pos[3 * p + ax];
                out_vel[3 * o + ax] = a.vel[3 * p + ax];
            }
            continue;
        }
        int i, j, k;
        reo_cell(a, c0 + l, i, j, k);
        float x[3];
        reseed_pos(a, c0 + l, i, j, k, sl - keep, x);
        // get_velocity_self(x); the subdomain's get_velocity samples at x - bmin
        const float qx = a.sub ? x[0] - a.bx : x[0], qy = a.sub ? x[1] - a.by : x[1],
                    qz = a.sub ? x[2] - a.bz : x[2];
        out_vel[3 * o] = sample_linear(gu, hd, qx, qy, qz);
        out_vel[3 * o + 1] = sample_linear(gv, hd, qx, qy, qz);
        out_vel[3 * o + 2] = sample_linear(gw, hd, qx, qy, qz);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) out_pos[3 * o + ax] = x[ax];
    }
}

// ------------------------------------------------------------------ helpers
static int ensure_capacity(vflip_particles* P, int np) {
    if ((size_t)np > P->cap) {
        const size_t cap = (size_t)np + (size_t)np / 4 + 1024;
        if (P->slot) VF_HIP(hipFree(P->slot));
        if (P->tmp) VF_HIP(hipFree(P->tmp));
        if (P->big) VF_HIP(hipFree(P->big));
        P->slot = P->tmp = P->big = nullptr;
        P->cap = 0;
        const size_t big_cap = cap / (kShortSegment + 1) + 2;
        if (hipMalloc(&P->slot, cap * sizeof(int)) != hipSuccess || hipMalloc(&P->tmp, cap * sizeof(int)) != hipSuccess ||
            hipMalloc(&P->big, (big_cap + 1) * sizeof(int)) != hipSuccess) {
            set_error("particle workspace of %zu entries: out of device memory", cap);
            return VFLIP_ERR_NOMEM;
        }
        P->cap = cap;
        P->big_cap = big_cap;
    }
    return 0;
}

size_t scan_chunk() { return kScanChunk; }

int ensure_selection_scratch(vflip_particles* P, size_t n) {
    if (n + 1 <= P->pflag_cap) return 0;
    if (P->pflag) VF_HIP(hipFree(P->pflag));
    if (P->ppartial) VF_HIP(hipFree(P->ppartial));
    P->pflag = P->ppartial = nullptr;
    P->pflag_cap = 0;
    const size_t cap = n + n / 4 + 1024;
    if (hipMalloc(&P->pflag, (cap + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&P->ppartial, (cap / kScanChunk + 1) * sizeof(int)) != hipSuccess) {
        set_error("selection workspace of %zu entries: out of device memory", cap);
        return VFLIP_ERR_NOMEM;
    }
    P->pflag_cap = cap;
    return 0;
}

void scan_exclusive_i32(int* a, size_t n, int* partial, hipStream_t s) {
    const unsigned nb = (unsigned)((n + kScanChunk - 1) / kScanChunk);
    if (nb == 0) {
        k_fill_i32<<<1, kBlock, 0, s>>>(a, 1, 0);
        return;
    }
    k_scan_reduce<<<nb, kBlock, 0, s>>>(n, a, partial);
    k_scan_partials<<<1, 1024, 0, s>>>((int)nb, partial, a + n);
    k_scan_apply<<<nb, kBlock, 0, s>>>(n, a, partial);
}

}  // namespace vflip

using namespace vflip;

// Particle level set of the liquid: at every cell centre (i h, j h, k h),
// phi = min |centre - p| over the particles binned in the 3x3x3 cells around it,
// minus radius (3h - radius where no particle is that close). One thread per
// cell gathers from vflip_hash's cell lists; the minimum does not depend on the
// visiting order, so the result is exact (float distances, no contraction).
namespace vflip {
// Tile of kLpX x kLpY x kLpZ cells per workgroup: the particles binned in the
// tile's halo (one cell around it) are staged in LDS once, then every cell
// scans its 3x3x3 neighbour cells from LDS. A minimum is independent of the
// order it is taken in, so the result equals the per-cell global scan bit for
// bit; tiles whose halo holds more than kLpCap particles scan global memory.
// The scans keep the least squared distance and take one square root at the
// end: a correctly rounded sqrtf is monotone, so sqrtf(min d^2) equals the
// minimum of the per-particle sqrtf(d^2).
constexpr int kLpX = 8, kLpY = 8, kLpZ = 4;
constexpr int kLpHX = kLpX + 2, kLpHY = kLpY + 2, kLpHZ = kLpZ + 2;
constexpr int kLpHalo = kLpHX * kLpHY * kLpHZ;  // 600
constexpr int kLpCap = 5400;  // 8 particles per cell over the whole halo is 4800
constexpr int kLpRead = 4;    // particles per step of the LDS scan
constexpr float kLpFar = 3.0e38f;  // pad-word coordinate: its squared distance is +inf

__device__ inline float cell_min2_global(int i, int j, int k, int ni, int nj, int nk, float cx, float cy, float cz,
                                         const float* __restrict__ pos, const int* __restrict__ cell_start,
                                         const int* __restrict__ order, float best2) {
    for (int ck = max(k - 1, 0); ck <= min(k + 1, nk - 1); ++ck)
        for (int cj = max(j - 1, 0); cj <= min(j + 1, nj - 1); ++cj)
            for (int ci = max(i - 1, 0); ci <= min(i + 1, ni - 1); ++ci) {
                const int q = (ck * nj + cj) * ni + ci;
                for (int s = cell_start[q]; s < cell_start[q + 1]; ++s) {
                    const int p = order[s];
                    const float dx = cx - pos[3 * p], dy = cy - pos[3 * p + 1], dz = cz - pos[3 * p + 2];
                    best2 = fminf(best2, dx * dx + dy * dy + dz * dz);
                }
            }
    return best2;
}

__global__ __launch_bounds__(kBlock) void k_liquid_phi(int ni, int nj, int nk, float h, float radius,
                                                       const float* __restrict__ pos,
                                                       const int* __restrict__ cell_start,
                                                       const int* __restrict__ order, float* __restrict__ phi,
                                                       int tiles_x, int tiles_y) {
    __shared__ int s_off[kLpHalo + 1];
    __shared__ int s_src[kLpHalo];  // cell_start of the halo cell minus its staged offset
    // halo cell e's particles sit at s_off[e] + e on: one pad word per cell, so
    // the segments of x-adjacent cells (adjacent lanes) start an odd stride apart
    // (about 9 words for 8 particles a cell) and hit different LDS banks
    // (plus kLpRead - 1 words that the unrolled scan may read past a run's end and masks)
    __shared__ float s_x[kLpCap + kLpHalo + kLpRead], s_y[kLpCap + kLpHalo + kLpRead],
        s_z[kLpCap + kLpHalo + kLpRead];
    const int t = threadIdx.x;
    const int tx = blockIdx.x % tiles_x, ty = (blockIdx.x / tiles_x) % tiles_y, tz = blockIdx.x / (tiles_x * tiles_y);
    const int x0 = tx * kLpX, y0 = ty * kLpY, z0 = tz * kLpZ;
    // halo cell e <-> (x0 - 1 + hx, y0 - 1 + hy, z0 - 1 + hz); cells outside the grid hold no particles
    auto halo_cell = [&](int e, int& q) {
        const int hx = e % kLpHX, hy = (e / kLpHX) % kLpHY, hz = e / (kLpHX * kLpHY);
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        const bool in = gx >= 0 && gy >= 0 && gz >= 0 && gx < ni && gy < nj && gz < nk;
        q = in ? (gz * nj + gy) * ni + gx : -1;
        return in;
    };
    // counts per halo cell, then an exclusive scan (thread t owns cells 3t..3t+2 of the 768 slots)
    int cnt[3], st[3], run = 0;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int e = 3 * t + u;
        int q;
        cnt[u] = 0;
        st[u] = 0;
        if (e < kLpHalo && halo_cell(e, q)) {
            st[u] = cell_start[q];
            cnt[u] = cell_start[q + 1] - st[u];
        }
        run += cnt[u];
    }
    int total;
    int before = block_exclusive_scan(run, total);
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int e = 3 * t + u;
        if (e < kLpHalo) {
            s_off[e] = before;
            s_src[e] = st[u] - before;
        }
        before += cnt[u];
        if (e < kLpHalo && total <= kLpCap) {
            s_x[before + e] = kLpFar;  // the pad word after cell e's particles
            s_y[before + e] = kLpFar;
            s_z[before + e] = kLpFar;
        }
    }
    if (t == 0) s_off[kLpHalo] = total;
    __syncthreads();
    const bool staged = total <= kLpCap;
    if (staged) {
        // one lane per staged particle: its halo cell is the last e with s_off[e] <= slot,
        // which is a non-empty cell inside the grid (cells outside hold no particles);
        // it lands at slot + e, after the pad words of cells 0..e-1
        // (four slots per pass, so each lane has four independent load chains in flight)
        constexpr int U = 4;
        for (int base = t; base < total; base += U * kBlock) {
            int p[U], dst[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int slot = base + u * kBlock;
                p[u] = -1;
                dst[u] = 0;
                if (slot < total) {
                    // fixed-length search (kLpHalo < 1024), unrolled so the U chains overlap
                    int lo = 0;
#pragma unroll
                    for (int step = 512; step > 0; step >>= 1)
                        if (lo + step < kLpHalo && s_off[lo + step] <= slot) lo += step;
                    p[u] = order[s_src[lo] + slot];
                    dst[u] = slot + lo;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (p[u] >= 0) {
                    s_x[dst[u]] = pos[3 * p[u]];
                    s_y[dst[u]] = pos[3 * p[u] + 1];
                    s_z[dst[u]] = pos[3 * p[u] + 2];
                }
            }
        }
    }
    __syncthreads();
    const int i = x0 + t % kLpX, j = y0 + (t / kLpX) % kLpY, k = z0 + t / (kLpX * kLpY);
    if (i >= ni || j >= nj || k >= nk) return;
    const float cx = (float)i * h, cy = (float)j * h, cz = (float)k * h;
    float best2 = INFINITY;
    if (staged) {
        // The three x-neighbour cells of a (y, z) neighbour row are consecutive
        // halo cells, so their particles and pad words form one run of LDS; runs
        // of rows outside the grid are empty. Each run is scanned kLpRead words a
        // step, the words past its end masked out.
        const int li = i - x0, lj = j - y0 + 1, lk = k - z0 + 1;  // li: halo x of the cell's left neighbour
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int e0 = ((lk + r / 3 - 1) * kLpHY + (lj + r % 3 - 1)) * kLpHX + li;
            const int a = s_off[e0] + e0, b = s_off[e0 + 3] + e0 + 3;
            for (int s = a; s < b; s += kLpRead) {
                float d2[kLpRead];
#pragma unroll
                for (int u = 0; u < kLpRead; ++u) {
                    const float dx = cx - s_x[s + u], dy = cy - s_y[s + u], dz = cz - s_z[s + u];
                    d2[u] = dx * dx + dy * dy + dz * dz;
                }
#pragma unroll
                for (int u = 0; u < kLpRead; ++u) best2 = fminf(best2, s + u < b ? d2[u] : INFINITY);
            }
        }
    } else {
        best2 = cell_min2_global(i, j, k, ni, nj, nk, cx, cy, cz, pos, cell_start, order, best2);
    }
    const float d = sqrtf(best2), cap = 3.0f * h;
    const float best = d < cap ? d : cap;
    phi[(k * nj + j) * ni + i] = best - radius;
}
}  // namespace vflip

extern "C" {

int vflip_particles_create(int ni, int nj, int nk, float h, vflip_particles** out) {
    VF_ARG(out, "out");
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && h > 0, "grid dims and spacing");
    VF_ARG((double)ni * nj * nk < 2147483647.0, "cell count must fit in int");
    auto* P = new vflip_particles();
    P->nx = ni;
    P->ny = nj;
    P->nz = nk;
    P->h = h;
    P->ncell = (size_t)ni * nj * nk;
    P->npartial = (P->ncell + kScanChunk - 1) / kScanChunk;
    if (hipMalloc(&P->partial, P->npartial * sizeof(int)) != hipSuccess) {
        delete P;
        set_error("scan partials: out of device memory");
        return VFLIP_ERR_NOMEM;
    }
    *out = P;
    return 0;
}

int vflip_particles_destroy(vflip_particles* P) {
    if (!P) return 0;
    if (P->slot) (void)hipFree(P->slot);
    if (P->tmp) (void)hipFree(P->tmp);
    if (P->big) (void)hipFree(P->big);
    if (P->partial) (void)hipFree(P->partial);
    if (P->reo) (void)hipFree(P->reo);
    if (P->pflag) (void)hipFree(P->pflag);
    if (P->ppartial) (void)hipFree(P->ppartial);
    delete P;
    return 0;
}

int vflip_particles_set_stream(vflip_particles* P, void* hip_stream) {
    VF_ARG(P, "handle");
    P->stream = (hipStream_t)hip_stream;
    return 0;
}

}  // extern "C"

namespace {
// The binning's launch sequence over caller-provided scratch: slot and tmp hold np
// ints, big 2 + np / (kShortSegment + 1), partial ceil(ncell / kScanChunk).
int hash_launches(hipStream_t s, const BinGeom& g, size_t nc, int np, const float* pos, int* cell_key,
                  int* cell_start, int* order, int* slot, int* tmp, int* big, int* partial, unsigned npartial) {
    k_fill_i32<<<min(grid_for(nc + 1), 4096u), kBlock, 0, s>>>(cell_start, nc + 1, 0);
    k_fill_i32<<<1, kBlock, 0, s>>>(big, 1, 0);
    if (np > 0) k_bin_count<<<grid_for(np), kBlock, 0, s>>>(np, pos, g, cell_key, slot, cell_start);
    k_scan_reduce<<<npartial, kBlock, 0, s>>>(nc, cell_start, partial);
    k_scan_partials<<<1, 1024, 0, s>>>((int)npartial, partial, cell_start + nc);
    k_scan_apply<<<npartial, kBlock, 0, s>>>(nc, cell_start, partial);
    if (np > 0) {
        k_bin_scatter<<<grid_for(np), kBlock, 0, s>>>(np, cell_key, slot, cell_start, order);
        k_cell_fix<<<grid_for(nc), kBlock, 0, s>>>(nc, cell_start, order, big);
        k_sort_long<<<256, kBlock, 0, s>>>(cell_start, order, tmp, big);
    }
    VF_CHECK_LAUNCH();
    return 0;
}

// Scratch of vflip_hash_noalloc: one array per device, as large as the stated limits
// need (about 34 MB); its device address is looked up once per device.
constexpr int kNoAllocMaxParticles = 1 << 22;
constexpr int kNoAllocMaxCells = 1 << 24;
constexpr size_t kNoAllocBig = (size_t)kNoAllocMaxParticles / (kShortSegment + 1) + 2;
constexpr size_t kNoAllocPartials = ((size_t)kNoAllocMaxCells + kScanChunk - 1) / kScanChunk;
constexpr size_t kNoAllocWords = 2 * (size_t)kNoAllocMaxParticles + kNoAllocBig + kNoAllocPartials;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) int scratch[kNoAllocWords];

namespace {
// The calling thread's current device's copy of `scratch` (looked up on its first use).
int scratch_ptr(int** out) {
    int dev = -1;
    VF_HIP(hipGetDevice(&dev));
    VF_ARG(dev >= 0 && dev < kMaxDevices, "device index");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        VF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (int*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int vflip_hash(vflip_particles* P, int np, const float* pos, int clamp, float bx, float by, float bz, int* cell_key,
               int* cell_start, int* order, int* nbinned) {
    VF_ARG(P && np >= 0 && cell_start, "hash args");
    VF_ARG(np == 0 || (pos && cell_key && order), "hash arrays");
    if (int rc = ensure_capacity(P, np > 0 ? np : 1)) return rc;
    hipStream_t s = P->stream;
    const size_t nc = P->ncell;
    const BinGeom g{P->h, P->nx, P->ny, P->nz, clamp ? 1 : 0, bx, by, bz};
    if (int rc = hash_launches(s, g, nc, np, pos, cell_key, cell_start, order, P->slot, P->tmp, P->big, P->partial,
                               (unsigned)P->npartial))
        return rc;
    if (nbinned) {
        VF_HIP(hipMemcpyAsync(nbinned, cell_start + nc, sizeof(int), hipMemcpyDeviceToHost, s));
        VF_HIP(hipStreamSynchronize(s));
    }
    return 0;
}

int vflip_hash_noalloc(int ni, int nj, int nk, float h, int np, const float* pos, int clamp, float bx, float by,
                       float bz, int* cell_key, int* cell_start, int* order, void* hip_stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && h > 0, "grid dims and spacing");
    VF_ARG(np >= 0 && np <= kNoAllocMaxParticles, "particle count beyond the no-alloc limit of 4194304");
    VF_ARG((double)ni * nj * nk <= (double)kNoAllocMaxCells, "cell count beyond the no-alloc limit of 16777216");
    VF_ARG(cell_start && (np == 0 || (pos && cell_key && order)), "hash arrays");
    int* base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    int* slot = base;
    int* tmp = slot + kNoAllocMaxParticles;
    int* big = tmp + kNoAllocMaxParticles;
    int* partial = big + kNoAllocBig;
    const size_t nc = (size_t)ni * nj * nk;
    const BinGeom g{h, ni, nj, nk, clamp ? 1 : 0, bx, by, bz};
    return hash_launches((hipStream_t)hip_stream, g, nc, np, pos, cell_key, cell_start, order, slot, tmp, big, partial,
                         (unsigned)((nc + kScanChunk - 1) / kScanChunk));
}

int vflip_p2g(vflip_particles* P, int variant, int method, int np, const float* pos, const float* vel,
              const int* cell_start, const int* order, float bx, float by, float bz, float* u, float* u_coef, float* v,
              float* v_coef, float* w, float* w_coef) {
    VF_ARG(P && np >= 0 && cell_start && (np == 0 || (pos && vel && order)), "p2g args");
    VF_ARG(variant == VFLIP_VARIANT_COARSE || variant == VFLIP_VARIANT_SUB, "variant");
    VF_ARG(method == 0 || method == 1 || method == 2, "method");
    VF_ARG((!u) == (!u_coef) && (!v) == (!v_coef) && (!w) == (!w_coef), "field and coef come in pairs");
    hipStream_t s = P->stream;
    const int nx = P->nx, ny = P->ny, nz = P->nz;
    const int sub = variant == VFLIP_VARIANT_SUB;
    // positions binned with bmin only in the subdomain form
    P2GArgs a{pos, vel, cell_start, order, nx, ny, nz, P->h, sub, sub ? bx : 0.f, sub ? by : 0.f, sub ? bz : 0.f,
              {u, v, w}, {u_coef, v_coef, w_coef}};
    if (method == 0) {
        const int tiles_x = (nx + kOT - 1) / kOT, tiles_y = (ny + kOT - 1) / kOT, tiles_z = (nz + kOT - 1) / kOT;
        const int ntiles = tiles_x * tiles_y * tiles_z;
        k_p2g_owner<<<ntiles, kBlock, 0, s>>>(a, tiles_x, tiles_y, ntiles);
    } else if (method == 2) {
        for (int c = 0; c < 3; ++c) {
            if (!a.field[c]) continue;
            const size_t n = (size_t)face_dim(c, 0, nx, ny, nz) * face_dim(c, 1, nx, ny, nz) * face_dim(c, 2, nx, ny, nz);
            k_fill_i32<<<min(grid_for(n), 8192u), kBlock, 0, s>>>((int*)a.field[c], n, 0);
            k_fill_i32<<<min(grid_for(n), 8192u), kBlock, 0, s>>>((int*)a.coef[c], n, 0);
        }
        const int tiles_x = (nx + kPTX - 1) / kPTX, tiles_y = (ny + kPTY - 1) / kPTY, tiles_z = (nz + kPTZ - 1) / kPTZ;
        const int ntiles = tiles_x * tiles_y * tiles_z;
        k_p2g_scatter<<<ntiles, kBlock, 0, s>>>(a, tiles_x, tiles_y, ntiles);
        for (int c = 0; c < 3; ++c) {
            if (!a.field[c]) continue;
            const size_t n = (size_t)face_dim(c, 0, nx, ny, nz) * face_dim(c, 1, nx, ny, nz) * face_dim(c, 2, nx, ny, nz);
            k_p2g_normalize<<<min(grid_for(n), 8192u), kBlock, 0, s>>>(c, nx, ny, nz, sub, a.field[c], a.coef[c]);
        }
    }
    if (method == 1 || sub) {
        for (int c = 0; c < 3; ++c) {
            if (!a.field[c]) continue;
            const int fx = face_dim(c, 0, nx, ny, nz), fy = face_dim(c, 1, nx, ny, nz), fz = face_dim(c, 2, nx, ny, nz);
            k_p2g_gather<<<dim3((fx + 63) / 64, (fy + 3) / 4, fz), dim3(64, 4), 0, s>>>(a, c, method == 1 ? 0 : 1);
        }
    }
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_liquid_phi(int ni, int nj, int nk, float h, float radius, const float* pos, const int* cell_start,
                     const int* order, float* phi, void* hip_stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && (size_t)ni * nj * nk < (1u << 31), "grid dims");
    VF_ARG(cell_start && order && phi, "level-set arrays");
    const int tiles_x = (ni + kLpX - 1) / kLpX, tiles_y = (nj + kLpY - 1) / kLpY, tiles_z = (nk + kLpZ - 1) / kLpZ;
    k_liquid_phi<<<tiles_x * tiles_y * tiles_z, kBlock, 0, (hipStream_t)hip_stream>>>(
        ni, nj, nk, h, radius, pos, cell_start, order, phi, tiles_x, tiles_y);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_get_delta(size_t n, const float* fnew, const float* fold, float* out, void* hip_stream) {
    VF_ARG(fnew && fold && out, "get_delta arrays");
    if (n == 0) return 0;
    k_get_delta<<<min(grid_for(n), 8192u), kBlock, 0, (hipStream_t)hip_stream>>>(n, fnew, fold, out);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_reorder(vflip_particles* P, int variant, float bx, float by, float bz, unsigned seed, int np,
                  const float* pos, const float* vel, const int* cell_start, const int* order, const float* u,
                  const float* v, const float* w, const float* solid_phi, int cap, float* out_pos, float* out_vel,
                  int* nout) {
    VF_ARG(P && nout && cell_start && u && v && w && solid_phi, "reorder args");
    VF_ARG(variant == VFLIP_VARIANT_COARSE || variant == VFLIP_VARIANT_SUB, "variant");
    VF_ARG(np >= 0 && cap >= 0, "reorder counts");
    VF_ARG(np == 0 || (pos && vel && order), "reorder particles");
    VF_ARG(cap == 0 || (out_pos && out_vel), "reorder outputs");
    hipStream_t s = P->stream;
    const size_t nc = P->ncell;
    if (!P->reo && hipMalloc(&P->reo, (2 * nc + 1) * sizeof(int)) != hipSuccess) {
        P->reo = nullptr;
        set_error("reorder workspace of %zu cells: out of device memory", nc);
        return VFLIP_ERR_NOMEM;
    }
    int* cnt = P->reo;
    int* mask = P->reo + nc + 1;
    const int sub = variant == VFLIP_VARIANT_SUB;
    const ReoArgs a{P->nx, P->ny, P->nz, P->h, sub, sub ? bx : 0.f, sub ? by : 0.f, sub ? bz : 0.f, seed,
                    pos, vel, cell_start, order, u, v, w, solid_phi};
    k_reorder_count<<<grid_for(nc, kReoTile), kBlock, 0, s>>>(nc, a, cnt, mask);
    const unsigned nb = (unsigned)P->npartial;
    k_scan_reduce<<<nb, kBlock, 0, s>>>(nc, cnt, P->partial);
    k_scan_partials<<<1, 1024, 0, s>>>((int)nb, P->partial, cnt + nc);
    k_scan_apply<<<nb, kBlock, 0, s>>>(nc, cnt, P->partial);
    VF_CHECK_LAUNCH();
    int total = 0;
    VF_HIP(hipMemcpyAsync(&total, cnt + nc, sizeof(int), hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    *nout = total;
    if (total > cap) {
        set_error("reorder: %d particles do not fit the output capacity %d", total, cap);
        return VFLIP_ERR_CAPACITY;
    }
    if (total > 0) k_reorder_emit<<<grid_for(nc, kReoTile), kBlock, 0, s>>>(nc, a, cnt, mask, out_pos, out_vel);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_g2p(int ni, int nj, int nk, float h, int np, const float* pos, float* vel, const float* u_new,
              const float* v_new, const float* w_new, const float* u_old, const float* v_old, const float* w_old,
              float flip_ratio, float bx, float by, float bz, void* hip_stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && h > 0 && np >= 0, "g2p dims");
    VF_ARG(np == 0 || (pos && vel), "g2p particles");
    VF_ARG(u_new && v_new && w_new, "g2p grids");
    VF_ARG((!u_old) == (!v_old) && (!v_old) == (!w_old), "old grids: all or none");
    VF_ARG(flip_ratio >= 1.0f || u_old, "a PIC blend (flip_ratio < 1) needs the old grids");
    if (np == 0) return 0;
    G2PArgs a{ni, nj, nk, h, {u_new, v_new, w_new}, {u_old, v_old, w_old}, flip_ratio, bx, by, bz};
    k_g2p<<<grid_for(np), kBlock, 0, (hipStream_t)hip_stream>>>(np, pos, vel, a);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_advect_particles(int ni, int nj, int nk, float h, float dt, int variant, float bx, float by, float bz, int np,
                           float* pos, const float* u, const float* v, const float* w, const float* solid_phi,
                           void* hip_stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && h > 0 && np >= 0, "advect_particles dims");
    VF_ARG(variant == VFLIP_VARIANT_COARSE || variant == VFLIP_VARIANT_SUB, "variant");
    VF_ARG(np == 0 || (pos && u && v && w && solid_phi), "advect_particles arrays");
    if (np == 0) return 0;
    AdvArgs a{ni, nj, nk, h, dt, variant == VFLIP_VARIANT_SUB, bx, by, bz, u, v, w, solid_phi};
    k_advect_particles<<<grid_for(np), kBlock, 0, (hipStream_t)hip_stream>>>(np, pos, a);
    VF_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
