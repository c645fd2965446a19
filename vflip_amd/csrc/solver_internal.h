// Internal solver state shared by the solver (mgpcg.hip) and the batched
// subdomain solver (batch.hip). Not part of the C ABI.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "mg_kernels.h"

using vflip::kZOut;
using vflip::kZTX;
using vflip::kVecBlocks;
using vflip::Op;
using vflip::Scalars;

struct LevelBuf {
    int ni = 0, nj = 0, nk = 0;
    size_t n = 0;
    float *d = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *leak = nullptr;
    __half *hd = nullptr, *hax = nullptr, *hay = nullptr, *haz = nullptr;
    float hscale = 1.f;
    bool half = false;
    float *x = nullptr, *xt = nullptr, *b = nullptr;
    void* vec_alloc = nullptr;
};

struct vflip_solver {
    int ni, nj, nk;
    size_t n;
    vflip_solve_opts opts;
    hipStream_t stream = nullptr;
    std::vector<LevelBuf> L;
    float *r = nullptr, *z = nullptr, *zt = nullptr, *s = nullptr, *q = nullptr, *xw = nullptr, *b32 = nullptr;
    void* pcg_alloc = nullptr;
    // FP64 iterate / level-0 diagonal / right-hand side (mixed-precision refinement)
    double *x64 = nullptr, *d64 = nullptr, *b64 = nullptr;
    bool d64_valid = false;  // d64 holds the assembled FP64 diagonal of the current operator
    bool b64_valid = false;  // b64 holds the assembled FP64 right-hand side
    double* partial = nullptr;
    int np_partial = 0;  // partial buffer capacity
    int np_last = 0;     // partials written by the last reduction
    Scalars* sc = nullptr;
    Scalars* h_sc = nullptr;  // pinned
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool setup_done = false;
    int zsweep_min_n = 32768;  // batched subdomain solves: levels this large use the fused sweep
    // single-grid solves, levels below 0: levels this large use the fused sweep,
    // smaller ones the per-colour sweeps (measured faster at 64^3 and below)
    int fused_min_n = 128 * 128 * 128;
    bool use_zrow = true;      // row-wave form of the fused sweep where the row length allows it
    int cus = 256;             // compute units of the device (launch sizing)
    int vblocks = kVecBlocks;  // workgroups (partials) of the grid-stride vector kernels
    // coarsest-level operator RBGS^coarse_sweeps tabulated as a dense matrix
    float* ctab = nullptr;
    float* cpart = nullptr;
    int ctab_sweeps = -1;
    bool use_ctab = true;
    // operator-change detection between assemblies
    int* d_changed = nullptr;
    int* h_changed = nullptr;
    bool check_changed = false;
    // level-0 smoother timing (profiling)
    bool prof = false;
    std::vector<hipEvent_t> prof_ev;
    int prof_used = 0;
    // The PCG iteration's two launch sequences (matvec + update, preconditioner
    // + direction update) captured as graphs, keyed by a hash of every argument
    // they bake in; captured on a private stream, launched on `stream`.
    bool use_graphs = false;  // measured slightly slower than direct launches on ROCm 7; VFLIP_GRAPHS=1
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t g_exec[2] = {nullptr, nullptr};
    unsigned long long g_key[2] = {0, 0};
    int g_np_last[2] = {0, 0};  // np_last after each sequence
    // FP64 pressure history of vflip_project_history: hist[hist_head] is the last
    // solution, hist[(hist_head + 2) % 3] the one before; hist_n of them are valid
    // solutions; a solve's result stays pending in x64 until the next guess commits it by
    // swapping buffers (x64 <-> the oldest slot). hist_G: Gram matrix p_a'A p_b over the
    // slots, FP64, one row added per history solve (by its final true residual).
    double* hist_alloc = nullptr;
    double* hist[3] = {nullptr, nullptr, nullptr};
    double* hist_part = nullptr;  // partial sums (Gram row, right-hand-side dots)
    int hist_part_n = 0;          // its capacity per sum
    double* hist_G = nullptr;     // 3 x 3
    double* hist_coef = nullptr;  // weights of p1, p2, p3
    int hist_n = 0, hist_head = 0;
    bool hist_pending = false;          // x64 holds a solution not yet committed to the history
    bool hist_solve = false;            // set during a history solve: the final true residual
    const double* gram_p[2] = {nullptr, nullptr};  // accumulates the Gram row against these
    int gram_nb = 0;                    // workgroups of that true-residual launch
    double* f64_alloc = nullptr;        // base of the x64 / d64 / b64 allocation
    // FP64 defect correction around FP32 solves (vflip_solver_solve_refined and the
    // batched form): the FP64 iterate, the FP32 residual and correction
    double* ref_x64 = nullptr;
    float *ref_r = nullptr, *ref_e = nullptr;
    void* ref_alloc = nullptr;
};

// Fused z-marching sweep geometry for a level. Rows of 128 or 256 cells use the
// row-wave kernel (one 1024-thread workgroup per compute unit, y tiles of kZOut
// rows, z chunks sized so the grid is one wave of workgroups); other shapes use
// the 64 x TY tile kernel.
struct ZGeom {
    int ty, tiles_x, tiles_y, zchunk, blocks;
    int v;  // cells per lane of the row-wave kernel, 0 for the tile kernel
};
inline ZGeom zgeom(const LevelBuf& lv, bool use_zrow, int cus) {
    ZGeom g;
    g.v = (use_zrow && (lv.ni == 128 || lv.ni == 256) && lv.nj >= 8) ? lv.ni / 64 : 0;
    if (g.v) {
        g.ty = kZOut;
        g.tiles_x = 1;
        g.tiles_y = (lv.nj + kZOut - 1) / kZOut;
        const int chunks = std::max(1, cus / g.tiles_y);
        g.zchunk = std::max(2, (lv.nk + chunks - 1) / chunks);
        g.blocks = g.tiles_y * ((lv.nk + g.zchunk - 1) / g.zchunk);
        return g;
    }
    g.ty = lv.nj >= 128 ? 8 : 4;
    if (const char* e = getenv("VFLIP_ZTY")) {
        const int t = atoi(e);
        if ((t == 4 || t == 8 || t == 16) && lv.nj >= t) g.ty = t;
    }
    g.tiles_x = (lv.ni + kZTX - 1) / kZTX;
    g.tiles_y = (lv.nj + g.ty - 1) / g.ty;
    const int tiles = g.tiles_x * g.tiles_y;
    // about 2048 workgroups (8 per CU), z chunks of at least 4 planes
    int zc = std::max(4, (int)((long long)lv.nk * tiles / 2048));
    const char* env = getenv("VFLIP_ZCHUNK");  // applies to the finest level only (tuning)
    if (env && atoi(env) > 0 && lv.nk >= 256) zc = atoi(env);
    zc = std::min(zc, lv.nk);
    g.zchunk = zc;
    g.blocks = tiles * ((lv.nk + zc - 1) / zc);
    return g;
}
inline ZGeom zgeom(const vflip_solver* S, const LevelBuf& lv) { return zgeom(lv, S->use_zrow, S->cus); }

// Levels at least this large use the fused z-marching sweep.
inline bool use_zsweep(const vflip_solver* S, const LevelBuf& lv, int min_n) {
    return (int)lv.n >= min_n && lv.ni >= 16 && lv.nj >= 4 && lv.nk >= 4;
}
inline bool use_zsweep(const vflip_solver* S, const LevelBuf& lv) { return use_zsweep(S, lv, S->zsweep_min_n); }

inline Op<float> make_op_f32(const LevelBuf& l) { return Op<float>{l.d, l.ax, l.ay, l.az, 1.f, l.ni, l.nj, l.nk}; }


// Outputs of an FP32 solve run by the batched subdomain solver (mgpcg.hip).
int solver_finish(vflip_solver* S, const float* rhs32, float* x32, int iterations, int status, vflip_solve_stats* st);
// FP64 defect correction rounds shared by vflip_solver_solve_refined and the batched
// form (mgpcg.hip): begin (x = 0, buffers, setup), residual into ref_r and sc->true_res,
// add the correction ref_e, end (outputs), stats.
int solver_refine_begin(vflip_solver* S, hipStream_t s, bool assembled);
int solver_refine_residual(vflip_solver* S, hipStream_t s, const float* rhs32, bool assembled);
int solver_refine_add(vflip_solver* S, hipStream_t s);
int solver_refine_end(vflip_solver* S, hipStream_t s, float* x32, double* x64);
void solver_refine_stats(int iterations, bool converged, double true_res, double bnorm, double ms,
                         vflip_solve_stats* st);
// The coarsest-level table for the current operator, built if missing or stale (mgpcg.hip).
int solver_ensure_ctab(vflip_solver* S);
