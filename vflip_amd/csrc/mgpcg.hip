// Multigrid-preconditioned CG pressure solver for gfx950 (host side).
//
// Restates the reference aggregation-AMG PCG (S/AlgebraicMultigrid.h:22-301,
// S/GeometricLevelGen.{h,cpp}) matrix-free: every level is a symmetric 7-point
// operator stored as diag + couplings to the +x/+y/+z neighbour (16 B/cell in
// FP32), built in closed form on the GPU instead of sparse triple products.
//
// Storage per level l (linear, x fastest):
//   d, ax, ay, az, leak  FP32 (leak = diag - sum|couplings|, the Dirichlet part,
//                        only read when building the next level)
//   hd, hax, hay, haz    FP16 copies scaled by 1/hscale (mixed precision, l >= 2)
//   x, xt, b             FP32 V-cycle vectors; x/xt ping-pong for the fused
//                        sweeps (level 0 uses the PCG vectors r and z/zt)
#include "common.h"

#include <chrono>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <cstdint>
#include <initializer_list>
#include <type_traits>

#include "mg_kernels.h"
#include "solver_internal.h"

using namespace vflip;

namespace {

int ensure_f64(vflip_solver* S);

template <typename CT> Op<CT> make_op(const LevelBuf& l);
template <> Op<float> make_op<float>(const LevelBuf& l) {
    return Op<float>{l.d, l.ax, l.ay, l.az, 1.f, l.ni, l.nj, l.nk};
}
template <> Op<__half> make_op<__half>(const LevelBuf& l) {
    return Op<__half>{l.hd, l.hax, l.hay, l.haz, l.hscale, l.ni, l.nj, l.nk};
}

int alloc(void** p, size_t bytes) {
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return VFLIP_ERR_NOMEM;
    }
    return 0;
}

bool mixed(const vflip_solver* S) { return S->opts.precision == VFLIP_PREC_MIXED; }

dim3 grid3(int ni, int nj, int nk) { return dim3((ni + 63) / 64, (nj + 3) / 4, nk); }
dim3 grid3z(int ni, int nj, int nk) { return dim3((ni + 63) / 64, (nj + 3) / 4, (nk + kMatvecKZ - 1) / kMatvecKZ); }
int nblocks(dim3 g) { return (int)(g.x * g.y * g.z); }


// Row-wave stencil geometry (matvec, residual restriction): y tiles of kSRows
// rows, z chunks (even, for the restriction) sized so the grid is one wave of
// workgroups, two per compute unit. v == 0: the level or a pointer does not fit
// the row kernel.
struct RowGeom {
    int v, tiles_y, zchunk, blocks;
};
bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }
RowGeom rowgeom(const vflip_solver* S, const LevelBuf& lv, std::initializer_list<const void*> ptrs) {
    RowGeom g{0, 0, 0, 0};
    if (!S->use_zrow || !(lv.ni == 128 || lv.ni == 256) || lv.nj < 8) return g;
    for (const void* p : ptrs)
        if (p && !aligned16(p)) return g;
    g.v = lv.ni / 64;
    g.tiles_y = (lv.nj + kSRows - 1) / kSRows;
    const int chunks = std::max(1, 2 * S->cus / g.tiles_y);
    int zc = (lv.nk + chunks - 1) / chunks;
    zc = std::max(2, zc + (zc & 1));
    g.zchunk = zc;
    g.blocks = g.tiles_y * ((lv.nk + zc - 1) / zc);
    return g;
}
template <int KIND, typename Acc, typename CT = float>
void launch_stencil_row(vflip_solver* S, const RowGeom& g, const LevelBuf& lv, const float* b, const float* s,
                        float* out, int ci, int cj, double* partial) {
    const Op<CT> op = make_op<CT>(lv);
    if (g.v == 4)
        k_stencil_row<4, KIND, Acc, CT><<<g.blocks, 64 * kSRows, 0, S->stream>>>(op, b, s, out, ci, cj, g.tiles_y,
                                                                           g.zchunk, partial);
    else
        k_stencil_row<2, KIND, Acc, CT><<<g.blocks, 64 * kSRows, 0, S->stream>>>(op, b, s, out, ci, cj, g.tiles_y,
                                                                           g.zchunk, partial);
}

template <int MODE, typename Acc, typename CT>
void launch_zsweep_mode(vflip_solver* S, const ZGeom& g, const Op<CT>& op, const float* b, const float* xin,
                        float* xout, const float* xc, int ci, int cj, double* dot) {
    if (g.v == 4) {
        k_rbgs_zrow<4, MODE, Acc, CT><<<g.blocks, 1024, 0, S->stream>>>(op, b, xin, xout, xc, ci, cj, g.tiles_y,
                                                                         g.zchunk, dot);
        return;
    }
    if (g.v == 2) {
        k_rbgs_zrow<2, MODE, Acc, CT><<<g.blocks, 1024, 0, S->stream>>>(op, b, xin, xout, xc, ci, cj, g.tiles_y,
                                                                         g.zchunk, dot);
        return;
    }
    if (g.ty == 16)
        k_rbgs_zsweep<16, MODE, Acc, CT><<<g.blocks, 256, 0, S->stream>>>(op, b, xin, xout, xc, ci, cj, g.tiles_x,
                                                                           g.tiles_y, g.zchunk, dot);
    else if (g.ty == 8)
        k_rbgs_zsweep<8, MODE, Acc, CT><<<g.blocks, 256, 0, S->stream>>>(op, b, xin, xout, xc, ci, cj, g.tiles_x,
                                                                          g.tiles_y, g.zchunk, dot);
    else
        k_rbgs_zsweep<4, MODE, Acc, CT><<<g.blocks, 256, 0, S->stream>>>(op, b, xin, xout, xc, ci, cj, g.tiles_x,
                                                                          g.tiles_y, g.zchunk, dot);
}

template <typename Acc, typename CT>
void launch_zsweep(vflip_solver* S, const LevelBuf& lv, const float* b, const float* xin, float* xout, int mode,
                   const float* xc, int ci, int cj, double* dot) {
    const ZGeom g = zgeom(lv, S->use_zrow, S->cus);
    const Op<CT> op = make_op<CT>(lv);
    if (mode == 1) launch_zsweep_mode<1, Acc, CT>(S, g, op, b, xin, xout, xc, ci, cj, dot);
    else if (mode == 2) launch_zsweep_mode<2, Acc, CT>(S, g, op, b, xin, xout, xc, ci, cj, dot);
    else launch_zsweep_mode<0, Acc, CT>(S, g, op, b, xin, xout, xc, ci, cj, dot);
}


// Pre- or post-smoothing of level l. With dot_done (level 0, post) the last
// fused sweep also leaves z.r partials in S->partial.
template <typename CT> int smooth_level(vflip_solver* S, int l, bool pre, const float* xc, bool* dot_done) {
    LevelBuf& lv = S->L[l];
    const float* b = (l == 0) ? S->r : lv.b;
    float** xp = (l == 0) ? &S->z : &lv.x;
    float** xtp = (l == 0) ? &S->zt : &lv.xt;
    const int sweeps = pre ? S->opts.pre_sweeps : S->opts.post_sweeps;
    const LevelBuf* c = pre ? nullptr : &S->L[l + 1];
    if (dot_done) *dot_done = false;
    if (sweeps == 0) {
        if (pre) {
            k_fill<<<grid_for(lv.n), kBlock, 0, S->stream>>>((unsigned)lv.n, *xp, 0.f);
        } else {
            k_prolong_add<<<grid_for(lv.n), kBlock, 0, S->stream>>>(lv.ni, lv.nj, lv.nk, *xp, xc, c->ni, c->nj);
        }
        VF_CHECK_LAUNCH();
        return 0;
    }
    // level 0 keeps the batched solver's rule (its last sweep fuses z.r, so the
    // summation order, and a batched solve's bits, stay those of the single solve)
    if (use_zsweep(S, lv, l == 0 ? S->zsweep_min_n : S->fused_min_n)) {
        for (int it = 0; it < sweeps; ++it) {
            const int mode = (it == 0) ? (pre ? 1 : 2) : 0;
            const bool last = (it == sweeps - 1);
            double* dot = (dot_done && last && l == 0) ? S->partial : nullptr;
            const bool prof = S->prof && l == 0 && mode == 0;  // the dominant sweep form
            if (prof) {
                while ((int)S->prof_ev.size() < S->prof_used + 2) {
                    hipEvent_t e;
                    VF_HIP(hipEventCreate(&e));
                    S->prof_ev.push_back(e);
                }
                VF_HIP(hipEventRecord(S->prof_ev[S->prof_used++], S->stream));
            }
            if (mixed(S))
                launch_zsweep<double, CT>(S, lv, b, *xp, *xtp, mode, xc, c ? c->ni : 0, c ? c->nj : 0, dot);
            else
                launch_zsweep<float, CT>(S, lv, b, *xp, *xtp, mode, xc, c ? c->ni : 0, c ? c->nj : 0, dot);
            VF_CHECK_LAUNCH();
            if (prof) VF_HIP(hipEventRecord(S->prof_ev[S->prof_used++], S->stream));
            std::swap(*xp, *xtp);
            if (dot) {
                S->np_last = zgeom(S, lv).blocks;
                *dot_done = true;
            }
        }
        return 0;
    }
    Op<CT> op = make_op<CT>(lv);
    float* x = *xp;
    const unsigned g = grid_for(lv.n);
    if (pre) {
        k_rbgs_first<CT><<<g, kBlock, 0, S->stream>>>(op, b, x);
    } else {
        k_rbgs_odd_prolong<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, xc, c->ni, c->nj);
    }
    k_rbgs_half<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, 0);
    for (int it = 1; it < sweeps; ++it) {
        k_rbgs_half<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, 1);
        k_rbgs_half<CT><<<g, kBlock, 0, S->stream>>>(op, b, x, 0);
    }
    VF_CHECK_LAUNCH();
    return 0;
}

template <typename CT> int restrict_level(vflip_solver* S, int l) {
    LevelBuf& lv = S->L[l];
    LevelBuf& c = S->L[l + 1];
    const float* b = (l == 0) ? S->r : lv.b;
    const float* x = (l == 0) ? S->z : lv.x;
    {
        const RowGeom rg = rowgeom(S, lv, {b, x, c.b});
        if (rg.v) {
            launch_stencil_row<1, float, CT>(S, rg, lv, b, x, c.b, c.ni, c.nj, nullptr);
            VF_CHECK_LAUNCH();
            return 0;
        }
    }
    const dim3 g((lv.ni + 63) / 64, (lv.nj + 1) / 2, (lv.nk + 1) / 2);
    k_residual_restrict<CT><<<g, dim3(64, 2, 2), 0, S->stream>>>(make_op<CT>(lv), b, x, c.b, c.ni, c.nj, c.nk);
    VF_CHECK_LAUNCH();
    return 0;
}

// z = M^-1 r, one V-cycle (S/AlgebraicMultigrid.h:130-188). When want_dot is set
// and the level-0 post-smoother fused it, partials of z.r are left in
// S->partial and *dot_done is set.
int run_vcycle(vflip_solver* S, bool want_dot, bool* dot_done) {
    const int nl = (int)S->L.size();
    int rc;
    if (dot_done) *dot_done = false;
    for (int l = 0; l < nl - 1; ++l) {
        rc = S->L[l].half ? smooth_level<__half>(S, l, true, nullptr, nullptr)
                          : smooth_level<float>(S, l, true, nullptr, nullptr);
        if (rc) return rc;
        rc = S->L[l].half ? restrict_level<__half>(S, l) : restrict_level<float>(S, l);
        if (rc) return rc;
    }
    {
        LevelBuf& lv = S->L[nl - 1];
        const float* b = (nl == 1) ? S->r : lv.b;
        float* x = (nl == 1) ? S->z : lv.x;
        const int ppr = (lv.ni + 1) / 2;
        if (lv.n > (size_t)kCoarseMax) {
            set_error("coarsest level %dx%dx%d exceeds the single-workgroup solver", lv.ni, lv.nj, lv.nk);
            return VFLIP_ERR_ARG;
        }
        if (S->ctab && S->ctab_sweeps == S->opts.coarse_sweeps) {
            const int nc = (int)lv.n;
            const int nch = (nc + kCoarseChunk - 1) / kCoarseChunk;
            k_coarse_apply_partial<<<dim3((nc + kBlock - 1) / kBlock, nch), kBlock, 0, S->stream>>>(S->ctab, b,
                                                                                                  S->cpart, nc);
            k_coarse_apply_reduce<<<(nc + kBlock - 1) / kBlock, kBlock, 0, S->stream>>>(S->cpart, x, nc, nch);
        } else if (!coarse_fits_registers(lv.ni, lv.nj, lv.nk)) {
            if (lv.half)
                k_coarsest_general<__half><<<1, kCoarseThreads, 0, S->stream>>>(make_op<__half>(lv), b, x,
                                                                                S->opts.coarse_sweeps, nullptr);
            else
                k_coarsest_general<float><<<1, kCoarseThreads, 0, S->stream>>>(make_op<float>(lv), b, x,
                                                                               S->opts.coarse_sweeps, nullptr);
        } else if (16 % ppr == 0) {  // x-neighbours through DPP (see k_coarsest)
            if (lv.half)
                k_coarsest<__half, true><<<1, kCoarseThreads, 0, S->stream>>>(make_op<__half>(lv), b, x,
                                                                              S->opts.coarse_sweeps, nullptr);
            else
                k_coarsest<float, true><<<1, kCoarseThreads, 0, S->stream>>>(make_op<float>(lv), b, x,
                                                                             S->opts.coarse_sweeps, nullptr);
        } else if (lv.half) {
            k_coarsest<__half><<<1, kCoarseThreads, 0, S->stream>>>(make_op<__half>(lv), b, x, S->opts.coarse_sweeps,
                                                                    nullptr);
        } else {
            k_coarsest<float><<<1, kCoarseThreads, 0, S->stream>>>(make_op<float>(lv), b, x, S->opts.coarse_sweeps,
                                                                   nullptr);
        }
        VF_CHECK_LAUNCH();
    }
    for (int l = nl - 2; l >= 0; --l) {
        const float* xc = S->L[l + 1].x;
        bool* dd = (l == 0 && want_dot) ? dot_done : nullptr;
        rc = S->L[l].half ? smooth_level<__half>(S, l, false, xc, dd) : smooth_level<float>(S, l, false, xc, dd);
        if (rc) return rc;
    }
    return 0;
}

// Tabulate the coarsest solve (coarse_sweeps RB-GS sweeps from a zero guess,
// a linear map of b): one workgroup per unit right-hand side.
int build_ctab(vflip_solver* S) {
    LevelBuf& lv = S->L.back();
    const int nc = (int)lv.n;
    if (!S->ctab) {
        int rc = alloc((void**)&S->ctab, (size_t)nc * nc * sizeof(float));
        if (!rc) rc = alloc((void**)&S->cpart, (size_t)nc * ((nc + kCoarseChunk - 1) / kCoarseChunk) * sizeof(float));
        if (rc) return rc;
    }
    if (lv.n > (size_t)kCoarseMax) {
        set_error("coarsest level %dx%dx%d exceeds the single-workgroup solver", lv.ni, lv.nj, lv.nk);
        return VFLIP_ERR_ARG;
    }
    const bool xreg = 16 % ((lv.ni + 1) / 2) == 0;  // x-neighbours through DPP (see k_coarsest)
    if (!coarse_fits_registers(lv.ni, lv.nj, lv.nk)) {
        if (lv.half)
            k_coarsest_general<__half><<<nc, kCoarseThreads, 0, S->stream>>>(make_op<__half>(lv), nullptr, nullptr,
                                                                             S->opts.coarse_sweeps, S->ctab);
        else
            k_coarsest_general<float><<<nc, kCoarseThreads, 0, S->stream>>>(make_op<float>(lv), nullptr, nullptr,
                                                                            S->opts.coarse_sweeps, S->ctab);
    } else if (lv.half && xreg)
        k_coarsest<__half, true><<<nc, kCoarseThreads, 0, S->stream>>>(make_op<__half>(lv), nullptr, nullptr,
                                                                       S->opts.coarse_sweeps, S->ctab);
    else if (lv.half)
        k_coarsest<__half><<<nc, kCoarseThreads, 0, S->stream>>>(make_op<__half>(lv), nullptr, nullptr,
                                                                 S->opts.coarse_sweeps, S->ctab);
    else if (xreg)
        k_coarsest<float, true><<<nc, kCoarseThreads, 0, S->stream>>>(make_op<float>(lv), nullptr, nullptr,
                                                                      S->opts.coarse_sweeps, S->ctab);
    else
        k_coarsest<float><<<nc, kCoarseThreads, 0, S->stream>>>(make_op<float>(lv), nullptr, nullptr,
                                                                S->opts.coarse_sweeps, S->ctab);
    VF_CHECK_LAUNCH();
    S->ctab_sweeps = S->opts.coarse_sweeps;
    return 0;
}

// o = (float)a over n cells, two per thread where the sizes and alignments allow.
void cast_f64_f32(unsigned n, const double* a, float* o, hipStream_t st) {
    if (n % 2 == 0 && ((uintptr_t)a & 15) == 0 && ((uintptr_t)o & 7) == 0)
        k_cast_f64_f32_x2<<<(n / 2 + kBlock - 1) / kBlock, kBlock, 0, st>>>(n, a, o);
    else
        k_cast_f64_f32<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(n, a, o);
}

int finalize(vflip_solver* S, int what, double tol = 0) {
    if (mixed(S))
        k_finalize<double><<<1, kBlock, 0, S->stream>>>(S->sc, S->partial, S->np_last, what, tol);
    else
        k_finalize<float><<<1, kBlock, 0, S->stream>>>(S->sc, S->partial, S->np_last, what, tol);
    VF_CHECK_LAUNCH();
    return 0;
}

int fetch_scalars(vflip_solver* S) {
    VF_HIP(hipMemcpyAsync(S->h_sc, S->sc, sizeof(Scalars), hipMemcpyDeviceToHost, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int check_dims(int ni, int nj, int nk) {
    if (ni < 1 || nj < 1 || nk < 1) return 0;
    if ((double)ni * nj * nk > 2147483647.0) return 0;
    return 1;
}

// Reductions record how many partials they wrote.
void launch_absmax(vflip_solver* S, unsigned n, const float* a) {
    k_absmax<<<S->vblocks, kBlock, 0, S->stream>>>(n, a, S->partial);
    S->np_last = S->vblocks;
}
void launch_dot(vflip_solver* S, unsigned n, const float* a, const float* b) {
    if (mixed(S)) k_dot<double><<<S->vblocks, kBlock, 0, S->stream>>>(n, a, b, S->partial);
    else k_dot<float><<<S->vblocks, kBlock, 0, S->stream>>>(n, a, b, S->partial);
    S->np_last = S->vblocks;
}
void launch_true_residual(vflip_solver* S, const float* b, const float* x, float* r) {
    const LevelBuf& l0 = S->L[0];
    dim3 g = grid3z(l0.ni, l0.nj, l0.nk);
    k_true_residual<float><<<g, dim3(64, 4), 0, S->stream>>>(make_op<float>(l0), b, x, r, S->partial);
    S->np_last = nblocks(g);
}
void launch_matvec_dot(vflip_solver* S, const float* s, float* q) {
    const LevelBuf& l0 = S->L[0];
    const RowGeom rg = rowgeom(S, l0, {s, q});
    if (rg.v) {
        if (mixed(S)) launch_stencil_row<0, double>(S, rg, l0, nullptr, s, q, 0, 0, S->partial);
        else launch_stencil_row<0, float>(S, rg, l0, nullptr, s, q, 0, 0, S->partial);
        S->np_last = rg.blocks;
        return;
    }
    dim3 g = grid3z(l0.ni, l0.nj, l0.nk);
    if (mixed(S))
        k_matvec_dot<float, double><<<g, dim3(64, 4), 0, S->stream>>>(make_op<float>(l0), s, q, S->partial);
    else
        k_matvec_dot<float, float><<<g, dim3(64, 4), 0, S->stream>>>(make_op<float>(l0), s, q, S->partial);
    S->np_last = nblocks(g);
}
void launch_update_xr(vflip_solver* S) {
    if (mixed(S))
        k_update_xr<double><<<S->vblocks, kBlock, 0, S->stream>>>((unsigned)S->n, S->sc, S->xw, S->s, S->r, S->q,
                                                                  S->partial);
    else
        k_update_xr<float><<<S->vblocks, kBlock, 0, S->stream>>>((unsigned)S->n, S->sc, S->xw, S->s, S->r, S->q,
                                                                 S->partial);
    S->np_last = S->vblocks;
}
void launch_update_s(vflip_solver* S) {
    if (mixed(S)) k_update_s<double><<<S->vblocks, kBlock, 0, S->stream>>>((unsigned)S->n, S->sc, S->s, S->z);
    else k_update_s<float><<<S->vblocks, kBlock, 0, S->stream>>>((unsigned)S->n, S->sc, S->s, S->z);
}

// z = M^-1 r followed by rho = z.r (fused into the last smoothing sweep when possible).
int precond_and_dot(vflip_solver* S, int what) {
    bool dot_done = false;
    int rc = run_vcycle(S, true, &dot_done);
    if (rc) return rc;
    if (!dot_done) launch_dot(S, (unsigned)S->n, S->z, S->r);
    VF_CHECK_LAUNCH();
    return finalize(S, what);
}

// FNV-1a over the host state the iteration's launches read.
struct KeyHash {
    unsigned long long h = 1469598103934665603ull;
    template <typename T> void add(const T& v) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
        for (size_t q = 0; q < sizeof(T); ++q) h = (h ^ p[q]) * 1099511628211ull;
    }
};

unsigned long long iteration_key(const vflip_solver* S, int which) {
    KeyHash k;
    k.add(which);
    k.add(S->stream);
    k.add(S->opts.pre_sweeps); k.add(S->opts.post_sweeps); k.add(S->opts.coarse_sweeps); k.add(S->opts.precision);
    k.add(S->use_zrow); k.add(S->zsweep_min_n); k.add(S->fused_min_n); k.add(S->vblocks); k.add(S->cus); k.add(S->n);
    k.add(S->r); k.add(S->z); k.add(S->zt); k.add(S->s); k.add(S->q); k.add(S->xw); k.add(S->x64);
    k.add(S->partial); k.add(S->sc);
    k.add(S->ctab); k.add(S->cpart); k.add(S->ctab_sweeps);
    for (const LevelBuf& l : S->L) {
        k.add(l.ni); k.add(l.nj); k.add(l.nk);
        k.add(l.d); k.add(l.ax); k.add(l.ay); k.add(l.az); k.add(l.leak);
        k.add(l.hd); k.add(l.hax); k.add(l.hay); k.add(l.haz); k.add(l.hscale); k.add(l.half);
        k.add(l.x); k.add(l.xt); k.add(l.b);
    }
    return k.h;
}

// The launches of one PCG iteration before the residual check (which == 0:
// q = A s with s.q, scalars, x and r updates with |r|, scalars) and after it
// (which == 1: z = M^-1 r with z.r, scalars, s = z + beta s).
int iteration_launches(vflip_solver* S, int which) {
    int rc;
    if (which == 0) {
        launch_matvec_dot(S, S->s, S->q);
        VF_CHECK_LAUNCH();
        if ((rc = finalize(S, 1))) return rc;
        if (mixed(S)) {
            k_update_x64r<<<S->vblocks, kBlock, 0, S->stream>>>((unsigned)S->n, S->sc, S->x64, S->s, S->r, S->q,
                                                                S->partial);
            S->np_last = S->vblocks;
        } else {
            launch_update_xr(S);
        }
        VF_CHECK_LAUNCH();
        return finalize(S, 2);
    }
    if ((rc = precond_and_dot(S, 3))) return rc;
    launch_update_s(S);
    VF_CHECK_LAUNCH();
    return 0;
}

// Runs an iteration sequence: replays its graph (capturing it first when the
// key changed), or launches it directly when graphs are off or profiling.
int run_iteration(vflip_solver* S, int which) {
    if (!S->use_graphs || S->prof) return iteration_launches(S, which);
    const unsigned long long key = iteration_key(S, which);
    if (!S->g_exec[which] || S->g_key[which] != key) {
        if (S->g_exec[which]) {
            VF_HIP(hipGraphExecDestroy(S->g_exec[which]));
            S->g_exec[which] = nullptr;
        }
        if (!S->cap_stream) VF_HIP(hipStreamCreateWithFlags(&S->cap_stream, hipStreamNonBlocking));
        hipStream_t user = S->stream;
        S->stream = S->cap_stream;
        VF_HIP(hipStreamBeginCapture(S->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = iteration_launches(S, which);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(S->cap_stream, &g);
        S->stream = user;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        VF_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&S->g_exec[which], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        VF_HIP(ei);
        S->g_key[which] = key;
        S->g_np_last[which] = S->np_last;
    }
    VF_HIP(hipGraphLaunch(S->g_exec[which], S->stream));
    S->np_last = S->g_np_last[which];
    return 0;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

void vflip_default_opts(vflip_solve_opts* o) {
    if (!o) return;
    o->tolerance = 1e-6;
    o->max_iterations = 50;
    o->precision = VFLIP_PREC_FP32;
    o->pre_sweeps = 4;
    o->post_sweeps = 4;
    o->coarse_sweeps = 200;
    // refreshing the FP64 true residual every 4 iterations keeps the FP32 recursive
    // residual's drift far below the tolerance: at 256^3 the mixed solve then stops
    // after 7 iterations instead of 9 (8 per refresh)
    o->refine_every = 4;
    o->warm_start = 0;
    o->half_from_level = 0;
}

int vflip_solver_destroy(vflip_solver* S) {
    if (!S) return 0;
    for (auto& l : S->L) {
        if (l.d) (void)hipFree(l.d);
        if (l.hd) (void)hipFree(l.hd);
        if (l.vec_alloc) (void)hipFree(l.vec_alloc);
    }
    if (S->pcg_alloc) (void)hipFree(S->pcg_alloc);
    if (S->f64_alloc) (void)hipFree(S->f64_alloc);
    if (S->hist_alloc) (void)hipFree(S->hist_alloc);
    if (S->ref_alloc) (void)hipFree(S->ref_alloc);
    if (S->ctab) (void)hipFree(S->ctab);
    if (S->cpart) (void)hipFree(S->cpart);
    if (S->d_changed) (void)hipFree(S->d_changed);
    if (S->h_changed) (void)hipHostFree(S->h_changed);
    for (auto e : S->prof_ev) (void)hipEventDestroy(e);
    for (auto g : S->g_exec)
        if (g) (void)hipGraphExecDestroy(g);
    if (S->cap_stream) (void)hipStreamDestroy(S->cap_stream);
    if (S->partial) (void)hipFree(S->partial);
    if (S->sc) (void)hipFree(S->sc);
    if (S->h_sc) (void)hipHostFree(S->h_sc);
    if (S->e0) (void)hipEventDestroy(S->e0);
    if (S->e1) (void)hipEventDestroy(S->e1);
    delete S;
    return 0;
}

int vflip_solver_set_opts(vflip_solver* S, const vflip_solve_opts* o) {
    VF_ARG(S && o, "solver/opts");
    VF_ARG(o->max_iterations >= 0 && o->pre_sweeps >= 0 && o->post_sweeps >= 0 && o->coarse_sweeps >= 0,
           "negative iteration or sweep count");
    VF_ARG(o->precision == VFLIP_PREC_FP32 || o->precision == VFLIP_PREC_MIXED, "precision");
    VF_ARG(o->half_from_level >= 0, "half_from_level");
    const bool need_half = (o->precision == VFLIP_PREC_MIXED);
    const bool had_half = (S->opts.precision == VFLIP_PREC_MIXED);
    const bool half_moved = o->half_from_level != S->opts.half_from_level;
    S->opts = *o;
    if (S->opts.refine_every < 0) S->opts.refine_every = 0;
    if (need_half != had_half || (need_half && half_moved)) S->setup_done = false;
    return 0;
}

int vflip_solver_create(int ni, int nj, int nk, const vflip_solve_opts* opts, vflip_solver** out) {
    VF_ARG(out, "out");
    *out = nullptr;
    VF_ARG(check_dims(ni, nj, nk), "grid dims");
    if (opts) {
        VF_ARG(opts->precision == VFLIP_PREC_FP32 || opts->precision == VFLIP_PREC_MIXED, "precision");
        VF_ARG(opts->half_from_level >= 0, "half_from_level");
        VF_ARG(opts->max_iterations >= 0 && opts->pre_sweeps >= 0 && opts->post_sweeps >= 0 &&
                   opts->coarse_sweeps >= 0,
               "negative iteration or sweep count");
    }
    vflip_solver* S = new vflip_solver();
    S->ni = ni; S->nj = nj; S->nk = nk;
    S->n = (size_t)ni * nj * nk;
    vflip_default_opts(&S->opts);
    if (opts) S->opts = *opts;
    if (const char* e = getenv("VFLIP_ZSWEEP_MIN_N")) S->zsweep_min_n = atoi(e);
    if (const char* e = getenv("VFLIP_FUSED_MIN_N")) S->fused_min_n = atoi(e);
    if (const char* e = getenv("VFLIP_ZROW")) S->use_zrow = atoi(e) != 0;
    if (const char* e = getenv("VFLIP_GRAPHS")) S->use_graphs = atoi(e) != 0;
    // grid-stride vector kernels: about 1024 cells per workgroup, at most kVecBlocks partials
    S->vblocks = (int)std::min<size_t>(kVecBlocks, std::max<size_t>(8, (S->n + 1023) / 1024));
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            S->cus = cus;
    }
    // level shapes, S/GeometricLevelGen.h:36-58
    {
        int a = ni, b = nj, c = nk;
        unsigned unknowns = (unsigned)ni * nj * nk;
        S->L.push_back(LevelBuf());
        S->L.back().ni = a; S->L.back().nj = b; S->L.back().nk = c;
        while (unknowns > 16u * 16u * 16u) {
            a = (int)std::ceil((float)a / 2.0);
            b = (int)std::ceil((float)b / 2.0);
            c = (int)std::ceil((float)c / 2.0);
            S->L.push_back(LevelBuf());
            S->L.back().ni = a; S->L.back().nj = b; S->L.back().nk = c;
            unknowns = (unsigned)a * b * c;
        }
    }
    int rc = 0;
    for (size_t l = 0; l < S->L.size() && !rc; ++l) {
        LevelBuf& lv = S->L[l];
        lv.n = (size_t)lv.ni * lv.nj * lv.nk;
        void* p = nullptr;
        rc = alloc(&p, 5 * lv.n * sizeof(float));
        if (rc) break;
        lv.d = (float*)p; lv.ax = lv.d + lv.n; lv.ay = lv.ax + lv.n; lv.az = lv.ay + lv.n; lv.leak = lv.az + lv.n;
        if (l >= 1) {
            rc = alloc(&lv.vec_alloc, 3 * lv.n * sizeof(float));
            if (rc) break;
            lv.x = (float*)lv.vec_alloc; lv.xt = lv.x + lv.n; lv.b = lv.xt + lv.n;
        }
        if (l >= 2) {
            rc = alloc(&p, 4 * lv.n * sizeof(__half));
            if (rc) break;
            lv.hd = (__half*)p; lv.hax = lv.hd + lv.n; lv.hay = lv.hax + lv.n; lv.haz = lv.hay + lv.n;
        }
    }
    if (!rc) {
        rc = alloc(&S->pcg_alloc, 7 * S->n * sizeof(float));
        if (!rc) {
            float* p = (float*)S->pcg_alloc;
            S->r = p; S->z = p + S->n; S->zt = p + 2 * S->n; S->s = p + 3 * S->n; S->q = p + 4 * S->n;
            S->xw = p + 5 * S->n; S->b32 = p + 6 * S->n;
        }
    }
    if (!rc) {
        const LevelBuf& l0 = S->L[0];
        int np = std::max(kVecBlocks, nblocks(grid3z(l0.ni, l0.nj, l0.nk)));
        // the true residual's max|r| and energy partials side by side
        np = std::max(np, 2 * nblocks(true_res_grid(l0.ni, l0.nj, l0.nk)));
        np = std::max(np, nblocks(dim3((l0.ni + 63) / 64, (l0.nj + 3) / 4, (l0.nk + kTrueResKZ - 1) / kTrueResKZ)));
        np = std::max(np, zgeom(l0, false, S->cus).blocks);
        np = std::max(np, zgeom(l0, true, S->cus).blocks);
        np = std::max(np, rowgeom(S, l0, {}).blocks);  // the row-wave matvec's s.q partials
        S->np_partial = np;
        rc = alloc((void**)&S->partial, (size_t)np * sizeof(double));
    }
    if (!rc) rc = alloc((void**)&S->sc, sizeof(Scalars));
    if (!rc) rc = alloc((void**)&S->d_changed, sizeof(int));
    if (!rc && hipHostMalloc((void**)&S->h_changed, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc failed");
        rc = VFLIP_ERR_NOMEM;
    }
    if (const char* e = getenv("VFLIP_COARSE_TABLE")) S->use_ctab = atoi(e) != 0;
    if (!rc) {
        if (hipHostMalloc((void**)&S->h_sc, sizeof(Scalars), hipHostMallocDefault) != hipSuccess) {
            set_error("hipHostMalloc failed");
            rc = VFLIP_ERR_NOMEM;
        }
    }
    if (!rc) {
        if (hipEventCreate(&S->e0) != hipSuccess || hipEventCreate(&S->e1) != hipSuccess) {
            set_error("hipEventCreate failed");
            rc = VFLIP_ERR_HIP;
        }
    }
    if (rc) {
        vflip_solver_destroy(S);
        return rc;
    }
    *out = S;
    return 0;
}

int vflip_solver_set_stream(vflip_solver* S, void* stream) {
    VF_ARG(S, "solver");
    S->stream = (hipStream_t)stream;
    return 0;
}

int vflip_solver_levels(const vflip_solver* S) { return S ? (int)S->L.size() : 0; }

int vflip_solver_level_dims(const vflip_solver* S, int l, int* ni, int* nj, int* nk) {
    VF_ARG(S && l >= 0 && l < (int)S->L.size(), "level");
    if (ni) *ni = S->L[l].ni;
    if (nj) *nj = S->L[l].nj;
    if (nk) *nk = S->L[l].nk;
    return 0;
}

int vflip_solver_set_matrix(vflip_solver* S, const float* d, const float* ax, const float* ay, const float* az,
                            const float* leak) {
    VF_ARG(S && d && ax && ay && az, "matrix arrays");
    LevelBuf& l0 = S->L[0];
    const size_t bytes = l0.n * sizeof(float);
    VF_HIP(hipMemcpyAsync(l0.d, d, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipMemcpyAsync(l0.ax, ax, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipMemcpyAsync(l0.ay, ay, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipMemcpyAsync(l0.az, az, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (leak) {
        VF_HIP(hipMemcpyAsync(l0.leak, leak, bytes, hipMemcpyDeviceToDevice, S->stream));
    } else {
        k_leak_from_diag<<<grid_for(l0.n), kBlock, 0, S->stream>>>(l0.ni, l0.nj, l0.nk, l0.d, l0.ax, l0.ay, l0.az,
                                                                   l0.leak);
        VF_CHECK_LAUNCH();
    }
    S->setup_done = false;
    S->d64_valid = false;
    S->b64_valid = false;
    return 0;
}

int vflip_solver_set_diag64(vflip_solver* S, const double* d64) {
    VF_ARG(S && d64, "solver/diag");
    int rc = ensure_f64(S);
    if (rc) return rc;
    VF_HIP(hipMemcpyAsync(S->d64, d64, S->n * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
    S->d64_valid = true;
    return 0;
}

int vflip_solver_setup(vflip_solver* S) {
    VF_ARG(S, "solver");
    if (S->check_changed && S->setup_done) {
        // the last assembly compared the new operator with the previous one
        VF_HIP(hipMemcpyAsync(S->h_changed, S->d_changed, sizeof(int), hipMemcpyDeviceToHost, S->stream));
        VF_HIP(hipStreamSynchronize(S->stream));
        S->check_changed = false;
        if (*S->h_changed == 0) {
            // same operator: the hierarchy stays valid, and the coarsest-level table
            // (dropped at the last change) now pays for itself
            if (S->use_ctab && S->ctab_sweeps != S->opts.coarse_sweeps) return build_ctab(S);
            return 0;
        }
    }
    const bool first = !S->setup_done;
    S->check_changed = false;
    // a new operator: the pressure history (and its Gram matrix) no longer applies
    S->hist_n = 0;
    S->hist_pending = false;
    VF_HIP(hipEventRecord(S->e0, S->stream));
    for (size_t l = 0; l + 1 < S->L.size(); ++l) {
        LevelBuf& f = S->L[l];
        LevelBuf& c = S->L[l + 1];
        k_coarsen<<<grid_for(c.n), kBlock, 0, S->stream>>>(f.ni, f.nj, f.nk, f.ax, f.ay, f.az, f.leak, c.ni, c.nj,
                                                          c.nk, c.d, c.ax, c.ay, c.az, c.leak);
        VF_CHECK_LAUNCH();
    }
    if (S->opts.precision == VFLIP_PREC_MIXED) {
        // FP16 copies of the levels from half_from_level on (default: every level) for
        // the V-cycle (the PCG matvec and the true residual keep the FP32 level 0).
        // The scale is the power of two that maps
        // the largest diagonal into (512, 1024], so integer coefficients (an
        // all-fluid interior's -1 couplings and 6 diagonals, and their Galerkin
        // coarsenings) stay exact.
        for (size_t l = 0; l < S->L.size(); ++l) {
            LevelBuf& lv = S->L[l];
            lv.half = false;
            if ((int)l < S->opts.half_from_level) continue;
            if (!lv.hd) {
                void* p = nullptr;
                int rc = alloc(&p, 4 * lv.n * sizeof(__half));
                if (rc) return rc;
                lv.hd = (__half*)p; lv.hax = lv.hd + lv.n; lv.hay = lv.hax + lv.n; lv.haz = lv.hay + lv.n;
            }
            launch_absmax(S, (unsigned)lv.n, lv.d);
            VF_CHECK_LAUNCH();
            int rc = finalize(S, 5);
            if (rc) return rc;
            rc = fetch_scalars(S);
            if (rc) return rc;
            const double mx = S->h_sc->true_res;
            const float scale = (mx > 0) ? (float)std::ldexp(1.0, (int)std::ceil(std::log2(mx / 1024.0))) : 1.f;
            lv.hscale = scale;
            const float inv = 1.f / scale;
            const unsigned g = grid_for(lv.n);
            k_to_half<<<g, kBlock, 0, S->stream>>>((unsigned)lv.n, lv.d, lv.hd, inv);
            k_to_half<<<g, kBlock, 0, S->stream>>>((unsigned)lv.n, lv.ax, lv.hax, inv);
            k_to_half<<<g, kBlock, 0, S->stream>>>((unsigned)lv.n, lv.ay, lv.hay, inv);
            k_to_half<<<g, kBlock, 0, S->stream>>>((unsigned)lv.n, lv.az, lv.haz, inv);
            VF_CHECK_LAUNCH();
            lv.half = true;
        }
    } else {
        for (auto& lv : S->L) lv.half = false;
    }
    // The coarsest-level table (below) costs about as much as a few solves'
    // worth of direct coarsest sweeps: build it for the first operator, and for
    // a changed one only once it is reused (an operator that changes every step,
    // like a liquid's, keeps the direct sweeps).
    S->ctab_sweeps = -1;
    if (S->use_ctab && first) {
        const int rc = build_ctab(S);
        if (rc) return rc;
    }
    VF_HIP(hipEventRecord(S->e1, S->stream));
    S->setup_done = true;
    return 0;
}

int vflip_solver_profile(vflip_solver* S, int enable) {
    VF_ARG(S, "solver");
    S->prof = enable != 0;
    S->prof_used = 0;
    return 0;
}

int vflip_solver_profile_read(vflip_solver* S, int* launches, double* total_ms, double* bytes_per_launch) {
    VF_ARG(S, "solver");
    VF_HIP(hipStreamSynchronize(S->stream));
    double tot = 0;
    for (int q = 0; q + 1 < S->prof_used; q += 2) {
        float ms = 0.f;
        VF_HIP(hipEventElapsedTime(&ms, S->prof_ev[q], S->prof_ev[q + 1]));
        tot += ms;
    }
    if (launches) *launches = S->prof_used / 2;
    if (total_ms) *total_ms = tot;
    // algorithmic bytes of one fused sweep on level 0: b and x_in read (4 B each), diag + 3
    // couplings read (4 B each in FP32, 2 B each from a mixed hierarchy's FP16 copy), x_out written
    const double coef = (!S->L.empty() && S->L[0].half) ? 2.0 : 4.0;
    if (bytes_per_launch) *bytes_per_launch = (12.0 + 4.0 * coef) * (double)S->n;
    S->prof_used = 0;
    return 0;
}

int vflip_solver_get_level(const vflip_solver* S, int l, float* d, float* ax, float* ay, float* az) {
    VF_ARG(S && l >= 0 && l < (int)S->L.size(), "level");
    const LevelBuf& lv = S->L[l];
    const size_t bytes = lv.n * sizeof(float);
    if (d) VF_HIP(hipMemcpyAsync(d, lv.d, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (ax) VF_HIP(hipMemcpyAsync(ax, lv.ax, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (ay) VF_HIP(hipMemcpyAsync(ay, lv.ay, bytes, hipMemcpyDeviceToDevice, S->stream));
    if (az) VF_HIP(hipMemcpyAsync(az, lv.az, bytes, hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int vflip_solver_vcycle(vflip_solver* S, const float* r, float* z) {
    VF_ARG(S && r && z, "solver/r/z");
    // an assembly since the last setup may have changed the operator (setup
    // compares it and keeps the hierarchy when it did not)
    if (!S->setup_done || S->check_changed) {
        int rc = vflip_solver_setup(S);
        if (rc) return rc;
    }
    VF_HIP(hipMemcpyAsync(S->r, r, S->n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
    int rc = run_vcycle(S, false, nullptr);
    if (rc) return rc;
    VF_HIP(hipMemcpyAsync(z, S->z, S->n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int vflip_solver_matvec(vflip_solver* S, const float* x, float* y) {
    VF_ARG(S && x && y, "solver/x/y");
    const LevelBuf& l0 = S->L[0];
    const RowGeom rg = rowgeom(S, l0, {x, y});
    if (rg.v)
        launch_stencil_row<0, float>(S, rg, l0, nullptr, x, y, 0, 0, nullptr);
    else
        k_matvec_dot<float, float><<<grid3z(l0.ni, l0.nj, l0.nk), dim3(64, 4), 0, S->stream>>>(make_op<float>(l0), x,
                                                                                               y, S->partial);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

}  // extern "C"

namespace {

int ensure_f64(vflip_solver* S) {
    if (S->f64_alloc) return 0;
    int rc = alloc((void**)&S->f64_alloc, 3 * S->n * sizeof(double));
    if (rc) return rc;
    S->x64 = S->f64_alloc;
    S->d64 = S->x64 + S->n;
    S->b64 = S->d64 + S->n;
    return 0;
}

// The FP64 true residual r = b - A x on level 0 (k_true_residual_row for rows of
// 128 or 256 cells, else k_true_residual64): max|r| partials in part, and
// optionally the energy partials after them and the Gram partials in gram.
// Returns the number of partials (workgroups).
int true_residual_pass(vflip_solver* S, hipStream_t st, const double* b64, const float* b32, const double* x, float* r,
                       double* part, double* gram, const double* p1, const double* p2, bool energy) {
    const LevelBuf& l0 = S->L[0];
    const double* d64 = S->d64_valid ? S->d64 : nullptr;
    if (l0.ni == 256 || l0.ni == 128) {
        constexpr int kZChunk = 16;
        const int tiles_y = (l0.nj + kTrRows - 1) / kTrRows, nzc = (l0.nk + kZChunk - 1) / kZChunk;
        const int nb = tiles_y * nzc;
        double* ep = energy ? part + nb : nullptr;
        const Op<float> op = make_op<float>(l0);
        if (l0.ni == 256) {
            if (gram) k_true_residual_row<4, true><<<nb, 64 * kTrRows, 0, st>>>(op, d64, b64, b32, x, r, part, tiles_y, kZChunk, gram, p1, p2, ep);
            else k_true_residual_row<4, false><<<nb, 64 * kTrRows, 0, st>>>(op, d64, b64, b32, x, r, part, tiles_y, kZChunk, nullptr, nullptr, nullptr, ep);
        } else {
            if (gram) k_true_residual_row<2, true><<<nb, 64 * kTrRows, 0, st>>>(op, d64, b64, b32, x, r, part, tiles_y, kZChunk, gram, p1, p2, ep);
            else k_true_residual_row<2, false><<<nb, 64 * kTrRows, 0, st>>>(op, d64, b64, b32, x, r, part, tiles_y, kZChunk, nullptr, nullptr, nullptr, ep);
        }
        return nb;
    }
    const dim3 g = true_res_grid(l0.ni, l0.nj, l0.nk);
    k_true_residual64<<<g, dim3(64, kTrueResTY), 0, st>>>(make_op<float>(l0), d64, b64, b32, x, r, part, gram, p1, p2,
                                                        energy ? part + nblocks(g) : nullptr);
    return nblocks(g);
}

// gram: also accumulate the history Gram row of x (vflip_project_history solves)
// energy: also the energy partials of x, in S->partial after the max|r| partials
// (finalize_energy reads them).
void launch_true_residual64(vflip_solver* S, const double* b64, const float* b32, const double* x, float* r,
                            bool gram = false, bool energy = false) {
    const int nb = true_residual_pass(S, S->stream, b64, b32, x, r, S->partial, gram ? S->hist_part : nullptr,
                                      gram ? S->gram_p[0] : nullptr, gram ? S->gram_p[1] : nullptr, energy);
    if (gram) S->gram_nb = nb;
    S->np_last = nb;
}

// PCG driver for both precisions (S/AlgebraicMultigrid.h:189-301).
//   plain : FP32 iterate; stop on the recursive |r|_inf as the reference.
//   mixed : FP64 iterate x64 (x += alpha s in FP64), FP64 dots and scalars; the
//           residual is replaced by the FP64 true residual b - A x (FP64 diagonal
//           and right-hand side when the assembly provided them) every
//           refine_every iterations and whenever the recursive residual meets
//           the tolerance; the stop test is on that true residual.
// rhs32 is required; rhs64 (optional) is the FP64 right-hand side.
int solve_impl(vflip_solver* S, const float* rhs32, const double* rhs64, float* x32, double* x64out,
               vflip_solve_stats* st) {
    int rc;
    float ms_setup = 0.f;
    // any other solve overwrites x64: a pending history entry (the last history solve's
    // result) is dropped, and with it the oldest entry, whose Gram row it took over
    if (!S->hist_solve && S->hist_pending) {
        S->hist_pending = false;
        S->hist_n = std::min(S->hist_n, 2);
    }
    // an assembly since the last setup may have changed the operator (setup
    // compares it and keeps the hierarchy when it did not)
    if (!S->setup_done || S->check_changed) {
        const bool fresh = !S->setup_done;
        rc = vflip_solver_setup(S);
        if (rc) return rc;
        if (fresh) {
            VF_HIP(hipEventSynchronize(S->e1));
            VF_HIP(hipEventElapsedTime(&ms_setup, S->e0, S->e1));
        }
    }
    if ((rc = ensure_f64(S))) return rc;
    const unsigned n = (unsigned)S->n;
    const bool mx = mixed(S);
    const unsigned gn = grid_for(n);
    VF_HIP(hipEventRecord(S->e0, S->stream));
    // |b|_inf and the stop threshold (S/AlgebraicMultigrid.h:212-233)
    if (rhs64) {
        k_absmax64<<<S->vblocks, kBlock, 0, S->stream>>>(n, rhs64, S->partial);
        S->np_last = S->vblocks;
    } else {
        launch_absmax(S, n, rhs32);
    }
    VF_CHECK_LAUNCH();
    if ((rc = finalize(S, 4, S->opts.tolerance))) return rc;
    // A mixed solve seeded with an FP32 guess vets it first: its energy from one FP32
    // stencil pass (k_energy_f32); only a guess that passes pays for the FP64 cast
    // and the FP64 true residual below.
    bool prevetted = false;
    if (S->opts.warm_start && mx && !x64out && x32) {
        const LevelBuf& l0 = S->L[0];
        const dim3 g((l0.ni + 63) / 64, (l0.nj + 3) / 4, (l0.nk + kTrueResKZ - 1) / kTrueResKZ);
        k_energy_f32<<<g, dim3(64, 4), 0, S->stream>>>(make_op<float>(l0), rhs64, rhs32, x32, S->partial);
        S->np_last = nblocks(g);
        VF_CHECK_LAUNCH();
        if ((rc = finalize(S, 6))) return rc;
        if ((rc = fetch_scalars(S))) return rc;
        prevetted = true;
    }
    if (S->opts.warm_start && prevetted && !(S->h_sc->energy < 0.0)) {
        // rejected: the zero start below (the same as after the FP64 vetting)
    } else if (S->opts.warm_start) {
        // initial guess from x32 (or x64out when given)
        if (x64out) {
            if (x64out != S->x64)
                VF_HIP(hipMemcpyAsync(S->x64, x64out, n * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
        } else {
            k_cast_f32_f64<<<gn, kBlock, 0, S->stream>>>(n, x32, S->x64);
        }
        if (!mx) cast_f64_f32(n, S->x64, S->xw, S->stream);
        // The PCG vets the guess: it minimises E(x) = x'Ax/2 - b'x, and
        // E(x0) < E(0) = 0 exactly when x0 is closer to the solution than zero
        // in the A-norm. A guess that is not is dropped (restart from zero). The
        // energy's sums come from the same pass as the residual.
        launch_true_residual64(S, mx ? rhs64 : nullptr, rhs32, S->x64, S->r, false, true);
        VF_CHECK_LAUNCH();
        const int ne = S->np_last;
        if ((rc = finalize(S, 2))) return rc;
        if (mx) k_finalize<double><<<1, kBlock, 0, S->stream>>>(S->sc, S->partial + ne, ne, 6, 0.0);
        else k_finalize<float><<<1, kBlock, 0, S->stream>>>(S->sc, S->partial + ne, ne, 6, 0.0);
        VF_CHECK_LAUNCH();
    } else {
        if (mx) k_fill64<<<gn, kBlock, 0, S->stream>>>(n, S->x64, 0.0);
        else k_fill<<<gn, kBlock, 0, S->stream>>>(n, S->xw, 0.f);
        k_copy<<<gn, kBlock, 0, S->stream>>>(n, S->r, rhs32);
        VF_CHECK_LAUNCH();
    }
    if ((rc = fetch_scalars(S))) return rc;
    vflip_solve_stats stats{};
    stats.ms_setup = ms_setup;
    if (S->opts.warm_start && !(S->h_sc->energy < 0.0)) {
        if (mx) k_fill64<<<gn, kBlock, 0, S->stream>>>(n, S->x64, 0.0);
        else k_fill<<<gn, kBlock, 0, S->stream>>>(n, S->xw, 0.f);
        k_copy<<<gn, kBlock, 0, S->stream>>>(n, S->r, rhs32);
        VF_CHECK_LAUNCH();
        S->h_sc->res = S->h_sc->bnorm;
        stats.warm_rejected = 1;
    }
    stats.guess_residual = S->h_sc->bnorm > 0 ? S->h_sc->res / S->h_sc->bnorm : 0.0;
    const double bnorm = S->h_sc->bnorm;
    const double tol_abs = S->h_sc->tol_abs;
    double res = S->h_sc->res;
    int status = VFLIP_OK;
    int it = 0;
    bool res_is_true = false;  // res is the FP64 true residual of the final iterate
    if (bnorm == 0) {
        if (mx) k_fill64<<<gn, kBlock, 0, S->stream>>>(n, S->x64, 0.0);
        else k_fill<<<gn, kBlock, 0, S->stream>>>(n, S->xw, 0.f);
        VF_CHECK_LAUNCH();
        stats.converged = 1;
        res = 0;
    } else if (res <= tol_abs && S->opts.warm_start) {
        stats.converged = 1;
    } else {
        if ((rc = precond_and_dot(S, 0))) return rc;
        if ((rc = fetch_scalars(S))) return rc;
        if (S->h_sc->breakdown) {
            status = VFLIP_ERR_BREAKDOWN;
        } else {
            k_copy<<<gn, kBlock, 0, S->stream>>>(n, S->s, S->z);
            VF_CHECK_LAUNCH();
            status = VFLIP_NOT_CONVERGED;
            for (it = 0; it < S->opts.max_iterations; ++it) {
                if ((rc = run_iteration(S, 0))) return rc;
                if ((rc = fetch_scalars(S))) return rc;
                res = S->h_sc->res;
                res_is_true = false;
                if (mx) {
                    const bool periodic = S->opts.refine_every > 0 && ((it + 1) % S->opts.refine_every) == 0;
                    if (periodic || res <= tol_abs) {
                        // residual replacement with the FP64 true residual
                        launch_true_residual64(S, rhs64, rhs32, S->x64, S->r, S->hist_solve);
                        VF_CHECK_LAUNCH();
                        if ((rc = finalize(S, 2))) return rc;
                        if ((rc = fetch_scalars(S))) return rc;
                        res = S->h_sc->res;
                        res_is_true = true;
                    }
                }
                if (res <= tol_abs) {
                    status = VFLIP_OK;
                    ++it;
                    break;
                }
                if ((rc = run_iteration(S, 1))) return rc;
            }
            stats.converged = (status == VFLIP_OK);
        }
    }
    // outputs
    if (mx) {
        if (x32) cast_f64_f32(n, S->x64, x32, S->stream);
        if (x64out && x64out != S->x64)
            VF_HIP(hipMemcpyAsync(x64out, S->x64, n * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
    } else {
        if (x32) VF_HIP(hipMemcpyAsync(x32, S->xw, n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
        k_cast_f32_f64<<<gn, kBlock, 0, S->stream>>>(n, S->xw, S->x64);
        if (x64out && x64out != S->x64)
            VF_HIP(hipMemcpyAsync(x64out, S->x64, n * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
    }
    VF_CHECK_LAUNCH();
    VF_HIP(hipEventRecord(S->e1, S->stream));
    // FP64 true residual of the returned solution (FP64 operator/rhs when available), for
    // the stats; a mixed solve that stopped on it already has it
    double true_res;
    if (res_is_true) {
        true_res = res;
        VF_HIP(hipEventSynchronize(S->e1));
    } else {
        launch_true_residual64(S, rhs64, rhs32, S->x64, nullptr, S->hist_solve);
        VF_CHECK_LAUNCH();
        if ((rc = finalize(S, 5))) return rc;
        if ((rc = fetch_scalars(S))) return rc;
        true_res = S->h_sc->true_res;
    }
    float ms = 0.f;
    VF_HIP(hipEventElapsedTime(&ms, S->e0, S->e1));
    stats.ms_solve = ms;
    stats.iterations = it;
    stats.residual = res;
    stats.rel_residual = bnorm > 0 ? res / bnorm : 0.0;
    stats.true_residual = bnorm > 0 ? true_res / bnorm : 0.0;
    if (st) *st = stats;
    return status;
}

}  // namespace

// The coarsest-level table for the current operator (built if missing or stale).
int solver_ensure_ctab(vflip_solver* S) {
    if (S->ctab && S->ctab_sweeps == S->opts.coarse_sweeps) return 0;
    return build_ctab(S);
}

// Outputs of an FP32 solve run by the batched subdomain solver (its PCG left
// x in S->xw and the scalars in S->sc): x32 = x, then the FP64 true residual
// and the stats, as vflip_solver_solve reports them.
int solver_finish(vflip_solver* S, const float* rhs32, float* x32, int iterations, int status,
                  vflip_solve_stats* st) {
    int rc;
    if ((rc = ensure_f64(S))) return rc;
    const unsigned n = (unsigned)S->n;
    if (x32) VF_HIP(hipMemcpyAsync(x32, S->xw, n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
    k_cast_f32_f64<<<grid_for(n), kBlock, 0, S->stream>>>(n, S->xw, S->x64);
    launch_true_residual64(S, nullptr, rhs32, S->x64, nullptr);
    VF_CHECK_LAUNCH();
    if ((rc = finalize(S, 5))) return rc;
    if ((rc = fetch_scalars(S))) return rc;
    if (st) {
        vflip_solve_stats stats{};
        const double bnorm = S->h_sc->bnorm;
        stats.iterations = iterations;
        stats.converged = status == VFLIP_OK;
        stats.residual = bnorm > 0 ? S->h_sc->res : 0.0;
        stats.rel_residual = bnorm > 0 ? S->h_sc->res / bnorm : 0.0;
        stats.true_residual = bnorm > 0 ? S->h_sc->true_res / bnorm : 0.0;
        *st = stats;
    }
    return 0;
}

// FP64 defect correction around FP32 solves (SubDomainSolver::pcg_projection solves in
// FP64 to 1e-9, S/sub_domain.cpp:739): the iterate is kept in FP64, each round solves
// A e = b - A x for the correction in FP32 and adds it. Buffers of the rounds; x = 0.
int solver_refine_begin(vflip_solver* S, hipStream_t s, bool assembled) {
    if (assembled && !S->b64_valid) {
        set_error("refined solve: the solver holds no assembled FP64 right-hand side");
        return VFLIP_ERR_ARG;
    }
    int rc;
    if (!S->setup_done || S->check_changed) {
        if ((rc = vflip_solver_setup(S))) return rc;
        VF_HIP(hipStreamSynchronize(S->stream));
    }
    if ((rc = ensure_f64(S))) return rc;
    if (!S->ref_alloc) {
        const size_t n = S->n;
        if ((rc = alloc(&S->ref_alloc, n * sizeof(double) + 2 * ((n + 63) & ~(size_t)63) * sizeof(float)))) return rc;
        S->ref_x64 = (double*)S->ref_alloc;
        S->ref_r = (float*)(S->ref_x64 + n);
        S->ref_e = S->ref_r + ((n + 63) & ~(size_t)63);
    }
    k_fill64<<<grid_for(S->n), kBlock, 0, s>>>((unsigned)S->n, S->ref_x64, 0.0);
    VF_CHECK_LAUNCH();
    return 0;
}
// ref_r = FP32(b - A x) in FP64 (b = the assembled FP64 right-hand side when `assembled`,
// else rhs32; the assembled FP64 diagonal when valid) and its max |.| into sc->true_res.
int solver_refine_residual(vflip_solver* S, hipStream_t s, const float* rhs32, bool assembled) {
    S->np_last = true_residual_pass(S, s, assembled ? S->b64 : nullptr, rhs32, S->ref_x64, S->ref_r, S->partial,
                                    nullptr, nullptr, nullptr, false);
    k_finalize<double><<<1, kBlock, 0, s>>>(S->sc, S->partial, S->np_last, 5, 0.0);
    VF_CHECK_LAUNCH();
    return 0;
}
// ref_x64 += ref_e
int solver_refine_add(vflip_solver* S, hipStream_t s) {
    k_add_f32_f64<<<grid_for(S->n), kBlock, 0, s>>>((unsigned)S->n, S->ref_e, S->ref_x64);
    VF_CHECK_LAUNCH();
    return 0;
}
// x32 = FP32(ref_x64), x64 (may be null) = ref_x64
int solver_refine_end(vflip_solver* S, hipStream_t s, float* x32, double* x64) {
    if (x32) cast_f64_f32((unsigned)S->n, S->ref_x64, x32, s);
    if (x64) VF_HIP(hipMemcpyAsync(x64, S->ref_x64, S->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    VF_CHECK_LAUNCH();
    return 0;
}
void solver_refine_stats(int iterations, bool converged, double true_res, double bnorm, double ms,
                         vflip_solve_stats* st) {
    if (!st) return;
    vflip_solve_stats o{};
    o.iterations = iterations;
    o.converged = converged ? 1 : 0;
    o.residual = true_res;
    o.rel_residual = bnorm > 0 ? true_res / bnorm : 0.0;
    o.true_residual = o.rel_residual;
    o.ms_solve = ms;
    o.guess_residual = 1.0;
    *st = o;
}

extern "C" {

int vflip_solver_solve(vflip_solver* S, const float* rhs, float* x, vflip_solve_stats* st) {
    VF_ARG(S && rhs && x, "solver/rhs/x");
    return solve_impl(S, rhs, nullptr, x, nullptr, st);
}

int vflip_solver_get_assembled64(vflip_solver* S, double* diag, double* rhs) {
    VF_ARG(S && (diag || rhs), "solver/outputs");
    VF_ARG(S->d64_valid && S->b64_valid, "the solver holds no assembled FP64 diagonal and right-hand side");
    if (diag) VF_HIP(hipMemcpyAsync(diag, S->d64, S->n * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
    if (rhs) VF_HIP(hipMemcpyAsync(rhs, S->b64, S->n * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int vflip_solver_solve_refined(vflip_solver* S, const float* rhs, float* x, double* x64, int assembled, double tol,
                               int max_rounds, vflip_solve_stats* st) {
    VF_ARG(S && rhs && (x || x64), "solver/rhs/x");
    VF_ARG(tol > 0 && max_rounds >= 0, "tolerance / rounds");
    const auto t0 = std::chrono::steady_clock::now();
    int rc = solver_refine_begin(S, S->stream, assembled != 0);
    if (rc) return rc;
    int its = 0;
    double bnorm = 0, tr = 0;
    bool ok = false;
    for (int round = 0;; ++round) {
        if ((rc = solver_refine_residual(S, S->stream, rhs, assembled != 0))) return rc;
        if ((rc = fetch_scalars(S))) return rc;
        tr = S->h_sc->true_res;
        if (round == 0) bnorm = tr;
        if (bnorm == 0 || tr <= tol * bnorm) {
            ok = true;
            break;
        }
        if (round == max_rounds) break;
        vflip_solve_stats in{};
        rc = solve_impl(S, S->ref_r, nullptr, S->ref_e, nullptr, &in);
        if (rc < 0) return rc;
        its += in.iterations;
        if ((rc = solver_refine_add(S, S->stream))) return rc;
    }
    if ((rc = solver_refine_end(S, S->stream, x, x64))) return rc;
    VF_HIP(hipStreamSynchronize(S->stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    solver_refine_stats(its, ok, tr, bnorm, ms, st);
    return ok ? VFLIP_OK : VFLIP_NOT_CONVERGED;
}

int vflip_solver_solve_f64(vflip_solver* S, const double* rhs, double* x, vflip_solve_stats* st) {
    VF_ARG(S && rhs && x, "solver/rhs/x");
    int rc = ensure_f64(S);
    if (rc) return rc;
    // FP32 copy of the right-hand side for the FP32 parts of the iteration
    cast_f64_f32((unsigned)S->n, rhs, S->b32, S->stream);
    VF_CHECK_LAUNCH();
    return solve_impl(S, S->b32, rhs, nullptr, x, st);
}

}  // extern "C"
namespace vflip {
int solver_history_guess(vflip_solver* S, const float* rhs32, double** guess, int* used);
}
extern "C" {

int vflip_solver_history_guess(vflip_solver* S, const float* rhs, double* guess) {
    VF_ARG(S && rhs && guess, "solver/rhs/guess");
    double* g = nullptr;
    int used = 0;
    int rc = solver_history_guess(S, rhs, &g, &used);
    if (rc) return rc;
    VF_HIP(hipMemcpyAsync(guess, g, S->n * sizeof(double), hipMemcpyDeviceToDevice, S->stream));
    VF_HIP(hipStreamSynchronize(S->stream));
    return used;
}

int vflip_solver_reset_history(vflip_solver* S) {
    VF_ARG(S, "solver");
    S->hist_n = 0;
    S->hist_pending = false;
    return 0;
}

int vflip_amgpcg_solve_host(int ni, int nj, int nk, const double* diag, const double* ax, const double* ay,
                            const double* az, const double* rhs, double* result, double tol, int maxit,
                            double* residual_out, int* iterations_out, const vflip_solve_opts* opts) {
    VF_ARG(diag && ax && ay && az && rhs && result, "host arrays");
    VF_ARG(check_dims(ni, nj, nk), "grid dims");
    vflip_solve_opts o;
    vflip_default_opts(&o);
    // The reference solves in FP64: without options the drop-in runs the mixed
    // solver, whose stop test is on the FP64 true residual, so residual_out is the
    // FP64 |b - A x|_inf the reference reports.
    o.precision = VFLIP_PREC_MIXED;
    if (opts) o = *opts;
    o.tolerance = tol;
    o.max_iterations = maxit;
    o.warm_start = 0;
    vflip_solver* S = nullptr;
    int rc = vflip_solver_create(ni, nj, nk, &o, &S);
    if (rc) return rc;
    const size_t n = (size_t)ni * nj * nk;
    std::vector<float> h(6 * n);
    float *hd = h.data(), *hx = hd + n, *hy = hx + n, *hz = hy + n, *hl = hz + n, *hb = hl + n;
    for (size_t c = 0; c < n; ++c) {
        hd[c] = (float)diag[c];
        hx[c] = (float)ax[c];
        hy[c] = (float)ay[c];
        hz[c] = (float)az[c];
        hb[c] = (float)rhs[c];
    }
    // leak in FP64: diag - sum of |couplings| of the row
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                const size_t c = lin3(i, j, k, ni, nj);
                double s = 0;
                if (i > 0) s -= ax[c - 1];
                if (j > 0) s -= ay[c - ni];
                if (k > 0) s -= az[c - (size_t)ni * nj];
                s -= ax[c] + ay[c] + az[c];
                const double l = diag[c] - s;
                hl[c] = (float)(l > 0 ? l : 0.0);
            }
    float* dev = nullptr;
    double* dev64 = nullptr;
    if (hipMalloc((void**)&dev, 6 * n * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&dev64, 3 * n * sizeof(double)) != hipSuccess) {
        if (dev) (void)hipFree(dev);
        vflip_solver_destroy(S);
        set_error("hipMalloc failed");
        return VFLIP_ERR_NOMEM;
    }
    vflip_solve_stats st{};
    int status = 0;
    do {
        // FP32 operator for the preconditioner and Krylov products; the exact FP64
        // diagonal and right-hand side for the true residual (mixed mode)
        if (hipMemcpy(dev, h.data(), 6 * n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dev64, diag, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dev64 + n, rhs, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("H2D copy failed");
            status = VFLIP_ERR_HIP;
            break;
        }
        status = vflip_solver_set_matrix(S, dev, dev + n, dev + 2 * n, dev + 3 * n, dev + 4 * n);
        if (status < 0) break;
        status = vflip_solver_set_diag64(S, dev64);
        if (status < 0) break;
        status = vflip_solver_solve_f64(S, dev64 + n, dev64 + 2 * n, &st);
        if (status < 0 && status != VFLIP_ERR_BREAKDOWN) break;
        if (hipMemcpy(result, dev64 + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("D2H copy failed");
            status = VFLIP_ERR_HIP;
            break;
        }
        if (residual_out) *residual_out = st.residual;
        if (iterations_out) *iterations_out = st.iterations;
    } while (0);
    (void)hipFree(dev);
    (void)hipFree(dev64);
    vflip_solver_destroy(S);
    return status;
}

}  // extern "C"

namespace vflip {
int solver_level0(vflip_solver* S, int* ni, int* nj, int* nk, float** d, float** ax, float** ay, float** az,
                  float** leak, hipStream_t* stream) {
    VF_ARG(S, "solver");
    LevelBuf& l = S->L[0];
    *ni = l.ni; *nj = l.nj; *nk = l.nk;
    *d = l.d; *ax = l.ax; *ay = l.ay; *az = l.az; *leak = l.leak;
    *stream = S->stream;
    return 0;
}
void solver_mark_matrix_changed(vflip_solver* S) { S->setup_done = false; }
// FP64 level-0 buffers for the assembly (diagonal, right-hand side) and the FP64 iterate.
int solver_f64(vflip_solver* S, double** d64, double** b64, double** x64) {
    int rc = ensure_f64(S);
    if (rc) return rc;
    *d64 = S->d64;
    *b64 = S->b64;
    *x64 = S->x64;
    return 0;
}
void solver_mark_f64_assembled(vflip_solver* S) {
    S->d64_valid = true;
    S->b64_valid = true;
}
bool solver_is_mixed(const vflip_solver* S) { return S->opts.precision == VFLIP_PREC_MIXED; }
// Returns the previous warm_start option and sets it to `on`.
int solver_set_warm_start(vflip_solver* S, int on) {
    const int old = S->opts.warm_start;
    S->opts.warm_start = on;
    return old;
}
// Solve the assembled system: FP32 rhs, plus the FP64 rhs in mixed mode.
int solver_solve_assembled(vflip_solver* S, const float* rhs32, float* x32, vflip_solve_stats* st) {
    const double* b64 = (S->opts.precision == VFLIP_PREC_MIXED && S->b64_valid) ? S->b64 : nullptr;
    return solve_impl(S, rhs32, b64, x32, nullptr, st);
}
// Before an assembly that rewrites level 0: clears the device change flag; the
// next setup only rebuilds the hierarchy when the assembly set it.
// The initial guess of vflip_project_history, written into the solver's FP64 iterate:
// zero with no history, else the energy-optimal combination of the last (up to
// three) pressures: f = P'b in one streaming pass, G from the Gram cache, the 3 x 3
// solve and the FP64 combination on the GPU, no host sync. A pending solution (the
// last solve's, still in x64) is committed first by swapping x64 with the oldest slot.
// rhs32: the assembled right-hand side (its FP64 form is used when valid).
// *used receives how many history entries the guess rests on.
int solver_history_guess(vflip_solver* S, const float* rhs32, double** guess, int* used) {
    int rc = ensure_f64(S);
    if (rc) return rc;
    if (!S->hist_alloc) {
        const LevelBuf& l0 = S->L[0];
        const int nb3 = nblocks(true_res_grid(l0.ni, l0.nj, l0.nk));
        S->hist_part_n = std::max(S->vblocks, nb3);
        const size_t words = 3 * S->n + 3 * (size_t)S->hist_part_n + 9 + 4;
        rc = alloc((void**)&S->hist_alloc, words * sizeof(double));
        if (rc) return rc;
        for (int q = 0; q < 3; ++q) S->hist[q] = S->hist_alloc + q * S->n;
        S->hist_part = S->hist_alloc + 3 * S->n;
        S->hist_G = S->hist_part + 3 * (size_t)S->hist_part_n;
        S->hist_coef = S->hist_G + 9;
        S->hist_n = 0;
        S->hist_pending = false;
    }
    if (S->hist_pending) {
        const int slot = (S->hist_head + 1) % 3;  // the oldest entry, dropped
        std::swap(S->x64, S->hist[slot]);
        S->hist_head = slot;
        S->hist_n = std::min(S->hist_n + 1, 3);
        S->hist_pending = false;
    }
    const unsigned n = (unsigned)S->n;
    double* g = S->x64;
    const int M = std::min(S->hist_n, 3);
    if (M == 0) {
        k_fill64<<<S->vblocks, kBlock, 0, S->stream>>>(n, g, 0.0);
    } else {
        const int a = S->hist_head, b = (S->hist_head + 2) % 3, c = (S->hist_head + 1) % 3;
        const double* b64 = (mixed(S) && S->b64_valid) ? S->b64 : nullptr;
        const double *p1 = S->hist[a], *p2 = M >= 2 ? S->hist[b] : nullptr, *p3 = M >= 3 ? S->hist[c] : nullptr;
        const int nb = S->vblocks;
        const bool pairs = n % 2 == 0 && ((uintptr_t)rhs32 & 7) == 0 && ((uintptr_t)b64 & 15) == 0 &&
                           ((uintptr_t)p1 & 15) == 0 && ((uintptr_t)p2 & 15) == 0 && ((uintptr_t)p3 & 15) == 0;
        if (pairs) {
            if (M == 1) k_hist_rhs_dots2<1><<<nb, kBlock, 0, S->stream>>>(n, b64, rhs32, p1, p2, p3, S->hist_part);
            else if (M == 2) k_hist_rhs_dots2<2><<<nb, kBlock, 0, S->stream>>>(n, b64, rhs32, p1, p2, p3, S->hist_part);
            else k_hist_rhs_dots2<3><<<nb, kBlock, 0, S->stream>>>(n, b64, rhs32, p1, p2, p3, S->hist_part);
        } else if (M == 1) {
            k_hist_rhs_dots<1><<<nb, kBlock, 0, S->stream>>>(n, b64, rhs32, p1, p2, p3, S->hist_part);
        } else if (M == 2) {
            k_hist_rhs_dots<2><<<nb, kBlock, 0, S->stream>>>(n, b64, rhs32, p1, p2, p3, S->hist_part);
        } else {
            k_hist_rhs_dots<3><<<nb, kBlock, 0, S->stream>>>(n, b64, rhs32, p1, p2, p3, S->hist_part);
        }
        if (M == 1) k_hist_solve<1><<<1, kBlock, 0, S->stream>>>(S->hist_part, nb, S->hist_G, a, b, c, S->hist_coef);
        else if (M == 2) k_hist_solve<2><<<1, kBlock, 0, S->stream>>>(S->hist_part, nb, S->hist_G, a, b, c, S->hist_coef);
        else k_hist_solve<3><<<1, kBlock, 0, S->stream>>>(S->hist_part, nb, S->hist_G, a, b, c, S->hist_coef);
        k_hist_combine<<<S->vblocks, kBlock, 0, S->stream>>>(n, S->hist_coef, p1, p2, p3, g);
    }
    VF_CHECK_LAUNCH();
    *guess = g;
    *used = M;
    return 0;
}
// r = rhs - A g for a guess (FP64 arithmetic with the assembled FP64 diagonal and,
// in mixed mode, right-hand side), stored FP32 in the solver's residual buffer.
int solver_guess_residual(vflip_solver* S, const float* rhs32, const double* g, float** r) {
    int rc = ensure_f64(S);
    if (rc) return rc;
    launch_true_residual64(S, (mixed(S) && S->b64_valid) ? S->b64 : nullptr, rhs32, g, S->r);
    VF_CHECK_LAUNCH();
    *r = S->r;
    return 0;
}
// The guess residual restricted by 4 per axis into the network's input grid
// (k_residual_restrict4): no full-resolution residual is stored.
int solver_guess_residual_restricted(vflip_solver* S, const float* rhs32, const double* g, float* crhs,
                                     float* csolid, float* cliquid) {
    int rc = ensure_f64(S);
    if (rc) return rc;
    const LevelBuf& l0 = S->L[0];
    if (l0.ni % 4 || l0.nj % 4 || l0.nk % 4) {
        set_error("restricted guess residual: grid %dx%dx%d not divisible by 4", l0.ni, l0.nj, l0.nk);
        return VFLIP_ERR_ARG;
    }
    const dim3 grid((l0.ni + kRR4X - 1) / kRR4X, l0.nj / 4, (l0.nk + kRR4Planes - 1) / kRR4Planes);
    k_residual_restrict4<<<grid, dim3(kRR4X, 4), 0, S->stream>>>(make_op<float>(l0), S->d64_valid ? S->d64 : nullptr,
                                                               (mixed(S) && S->b64_valid) ? S->b64 : nullptr, rhs32,
                                                               g, crhs, csolid, cliquid);
    VF_CHECK_LAUNCH();
    return 0;
}
// FP32 scratch of the solver's size that is free until the next solve.
float* solver_scratch32(vflip_solver* S) { return S->z; }
// guess += correction (FP32)
int solver_add_correction(vflip_solver* S, double* g, const float* corr) {
    k_add_f32_f64<<<S->vblocks, kBlock, 0, S->stream>>>((unsigned)S->n, corr, g);
    VF_CHECK_LAUNCH();
    return 0;
}
// Solve the assembled system from the guess in x64 (warm start; a guess that is no
// better than zero in the A-norm is dropped by the solve). The solution stays in x64,
// pending, and its final true residual adds its Gram row (against the two newest
// committed entries) in the slot that the next guess commits it to. x32 receives
// the FP32 pressure.
int solver_solve_history(vflip_solver* S, const float* rhs32, double* g, float* x32, vflip_solve_stats* st,
                         bool warm) {
    const double* b64 = (mixed(S) && S->b64_valid) ? S->b64 : nullptr;
    const int old = S->opts.warm_start;
    const int slot = (S->hist_head + 1) % 3;
    const int M = std::min(S->hist_n, 3);
    const int a = M >= 1 ? S->hist_head : -1, b = M >= 2 ? (S->hist_head + 2) % 3 : -1;
    S->opts.warm_start = warm ? 1 : 0;
    S->hist_solve = true;
    S->gram_p[0] = a >= 0 ? S->hist[a] : nullptr;
    S->gram_p[1] = b >= 0 ? S->hist[b] : nullptr;
    S->gram_nb = 0;
    const int status = solve_impl(S, rhs32, b64, x32, g, st);
    S->hist_solve = false;
    S->gram_p[0] = S->gram_p[1] = nullptr;
    S->opts.warm_start = old;
    if (status < 0 && status != VFLIP_ERR_BREAKDOWN) return status;
    if (status == VFLIP_OK && S->gram_nb > 0) {
        k_hist_gram_commit<<<1, kBlock, 0, S->stream>>>(S->hist_part, S->gram_nb, slot, a, b, S->hist_G);
        VF_CHECK_LAUNCH();
        S->hist_pending = true;
    } else {
        S->hist_n = 0;  // no trusted history after a failed solve
        S->hist_pending = false;
    }
    return status;
}

// A flag that an earlier assembly set and no setup has consumed yet stays set:
// assemblies in a row accumulate into it.
int* solver_begin_assembly(vflip_solver* S) {
    int* f = S->d_changed;
    if (!S->check_changed)
        k_fill<<<1, kBlock, 0, S->stream>>>(1, (float*)f, 0.f);  // 0.0f and int 0 share the bit pattern
    S->check_changed = true;
    return f;
}
}  // namespace vflip
