// Batched MGPCG for many independent subdomain systems (the nested fine grids).
//
// Every subdomain keeps its own solver (vflip_solver: hierarchy, vectors,
// scalars). A batch replays the solver's V-cycle / PCG schedule for all of them
// at once: each step is ONE launch whose blocks are split among the subdomains
// (a block finds its subdomain from the per-item first-block table and runs the
// same kernel body as the single-grid solver, with that subdomain's own block
// coordinates). The CG scalars never leave the GPU; a per-subdomain state word
// (done / iterations / status) is updated by the finalize step, and the blocks of
// a finished subdomain return at once. That makes a fixed number of PCG
// iterations a static launch sequence, captured once into a hipGraph and
// replayed until every subdomain has converged (one host check per replay).
//
// Levels of different subdomains line up by depth: all pre-smoothing and
// restrictions from the finest level down, every coarsest-level solve in one
// step, then the post-smoothing back up. Within a subdomain the order of
// operations is exactly the single-grid solver's, so with the same tile
// geometry a batched solve reproduces vflip_solver_solve bit for bit. An
// autotuner times the fused sweep's tile shapes (rows per tile, z chunk) for each
// level depth over the whole batch and keeps the fastest.
#include "common.h"

#include <algorithm>
#include <chrono>
#include <map>
#include <tuple>
#include <vector>

#include "mg_kernels.h"
#include "solver_internal.h"

using namespace vflip;

namespace {

struct BState {
    int done, iters, status, maxit;
};

// One subdomain's part of a batched launch; the meaning of the generic fields
// is given per kernel below.
struct BItem {
    int begin;            // first block of the item in the launch
    unsigned gx, gy, gz;  // the item's own grid
    int slot;             // subdomain index
    Op<float> op;
    const float *a0, *a1, *a2;
    float *o0, *o1;
    int i0, i1, i2, i3, i4;
    unsigned n;
    double* partial;
    Scalars* sc;
    double tol;
};

__device__ inline int bfind(const BItem* it, int n) {
    int q = 0;
    while (q + 1 < n && it[q + 1].begin <= (int)blockIdx.x) ++q;
    return q;
}
__device__ inline Blk bblk(const BItem& I) {
    const unsigned l = blockIdx.x - (unsigned)I.begin;
    return Blk{l % I.gx, (l / I.gx) % I.gy, l / (I.gx * I.gy), I.gx, I.gy, I.gz};
}

// sweep: a0 b, a1 x_in, o0 x_out, a2 xc, i0..i4 ci cj tiles_x tiles_y zchunk, partial: z.r partials or null
template <int TY, int MODE>
__global__ __launch_bounds__(256) void kb_zsweep(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    rbgs_zsweep_body<TY, MODE, float, float>(I.op, I.a0, I.a1, I.o0, I.a2, I.i0, I.i1, I.i2, I.i3, I.i4, I.partial, bblk(I));
}
// flat half sweeps: a0 b, o0 x (in place), a2 xc, i0 ci, i1 cj
__global__ __launch_bounds__(kBlock) void kb_first(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    rbgs_first_body<float>(I.op, I.a0, I.o0, bblk(I));
}
__global__ __launch_bounds__(kBlock) void kb_half(const BItem* __restrict__ it, int n, const BState* __restrict__ st,
                                                  int colour) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    rbgs_half_body<float>(I.op, I.a0, I.o0, colour, bblk(I));
}
__global__ __launch_bounds__(kBlock) void kb_odd_prolong(const BItem* __restrict__ it, int n,
                                                         const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    rbgs_odd_prolong_body<float>(I.op, I.a0, I.o0, I.a2, I.i0, I.i1, bblk(I));
}
// restriction (64x2x2 blocks): a0 b, a1 x, o0 b_coarse, i0..i2 coarse dims
__global__ __launch_bounds__(256) void kb_restrict(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    residual_restrict_body<float>(I.op, I.a0, I.a1, I.o0, I.i0, I.i1, I.i2, bblk(I));
}
// coarsest level, tabulated: a0 table, a1 b, o0 chunk partials, i0 cells / a0 partials, o0 x, i0 cells, i1 chunks
__global__ __launch_bounds__(kBlock) void kb_cpart(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    coarse_apply_partial_body(I.a0, I.a1, I.o0, I.i0, bblk(I));
}
__global__ __launch_bounds__(kBlock) void kb_cred(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    coarse_apply_reduce_body(I.a0, I.o0, I.i0, I.i1, bblk(I));
}
// q = A s and s.q partials (64x4 blocks): a0 s, o0 q
__global__ __launch_bounds__(kBlock) void kb_matvec(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    matvec_dot_body<float, float>(I.op, I.a0, I.o0, I.partial, bblk(I));
}
// vector kernels over n cells
__global__ __launch_bounds__(kBlock) void kb_dot(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    dot_body<float>(I.n, I.a0, I.a1, I.partial, bblk(I));
}
__global__ __launch_bounds__(kBlock) void kb_absmax(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    absmax_body(I.n, I.a0, I.partial, bblk(I));
}
// x += alpha s, r -= alpha q: o0 x, a0 s, o1 r, a1 q
__global__ __launch_bounds__(kBlock) void kb_update_xr(const BItem* __restrict__ it, int n,
                                                       const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    update_xr_body<float>(I.n, I.sc, I.o0, I.a0, I.o1, I.a1, I.partial, bblk(I));
}
// s = z + beta s: o0 s, a0 z
__global__ __launch_bounds__(kBlock) void kb_update_s(const BItem* __restrict__ it, int n,
                                                      const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    update_s_body<float>(I.n, I.sc, I.o0, I.a0, bblk(I));
}
__global__ __launch_bounds__(kBlock) void kb_fill(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    fill_body(I.n, I.o0, 0.f, bblk(I));
}
__global__ __launch_bounds__(kBlock) void kb_copy(const BItem* __restrict__ it, int n, const BState* __restrict__ st) {
    const BItem I = it[bfind(it, n)];
    if (st[I.slot].done) return;
    copy_body(I.n, I.o0, I.a0, bblk(I));
}
// CG scalars, one block per subdomain: i0 partial count, i1 what (k_finalize's codes), tol.
// Also the stop test the single-grid solver makes on the host: |b| == 0, breakdown
// after the first preconditioner apply, |r| <= tol |b| or the iteration limit.
__global__ __launch_bounds__(kBlock) void kb_finalize(const BItem* __restrict__ it, BState* __restrict__ st) {
    const BItem I = it[blockIdx.x];
    BState& S = st[I.slot];
    if (S.done) return;
    finalize_body<float>(I.sc, I.partial, I.i0, I.i1, I.tol);
    if (threadIdx.x != 0) return;
    const int what = I.i1;
    if (what == 4 && I.sc->bnorm == 0) {
        S.done = 1;
        S.status = VFLIP_OK;
    } else if (what == 0 && I.sc->breakdown) {
        S.done = 1;
        S.status = VFLIP_ERR_BREAKDOWN;
    } else if (what == 2) {
        S.iters += 1;
        if (I.sc->res <= I.sc->tol_abs) {
            S.done = 1;
            S.status = VFLIP_OK;
        } else if (S.iters >= S.maxit) {
            S.done = 1;
            S.status = VFLIP_NOT_CONVERGED;
        }
    }
}
__global__ void kb_reset(BState* st, const int* maxit, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) st[q] = BState{0, 0, VFLIP_NOT_CONVERGED, maxit[q]};
}

enum Kind { K_ZSWEEP, K_FIRST, K_HALF, K_ODD, K_RESTRICT, K_CPART, K_CRED, K_MATVEC, K_DOT, K_ABSMAX, K_UPXR, K_UPS,
            K_FILL, K_COPY, K_FIN };

struct Phase {
    Kind kind;
    int ty = 0, mode = 0, colour = 0;
    dim3 block{kBlock, 1, 1};
    std::vector<BItem> items;
    size_t off = 0;  // index of items[0] in the device item array
    int blocks = 0;
};

// Tile geometry of the fused sweep (the single-grid solver's unless tuned).
struct Tune {
    int ty = 0, zc = 0;
};

}  // namespace

struct vflip_batch {
    std::vector<vflip_solver*> S;
    hipStream_t stream = nullptr;
    std::vector<Tune> tune;  // per level depth
    std::vector<Phase> init, iter;
    BItem* d_items = nullptr;
    size_t d_items_cap = 0;
    BState* d_st = nullptr;
    BState* h_st = nullptr;
    int* d_maxit = nullptr;
    bool planned = false;
    int graph_iters = 8;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

ZGeom tile_geom(const vflip_batch* B, const vflip_solver* S, int l) {
    const LevelBuf& lv = S->L[l];
    ZGeom g = zgeom(lv, false, S->cus);
    if (l < (int)B->tune.size() && B->tune[l].ty) {
        g.ty = B->tune[l].ty;
        g.tiles_y = (lv.nj + g.ty - 1) / g.ty;
        g.zchunk = std::min(B->tune[l].zc, lv.nk);
        g.blocks = g.tiles_x * g.tiles_y * ((lv.nk + g.zchunk - 1) / g.zchunk);
    }
    return g;
}

// Builds phases step by step; items of the same step and launch shape share a phase.
struct Planner {
    vflip_batch* B;
    std::vector<Phase>* out;
    std::map<std::tuple<int, int, int, int>, size_t> open;  // (kind, ty, mode, colour) -> phase in this step
    // per subdomain: the current V-cycle vectors (x / x-tilde per level; level 0: z / zt)
    std::vector<std::vector<float*>> x, xt;

    void step() { open.clear(); }
    Phase& get(Kind k, int ty = 0, int mode = 0, int colour = 0, dim3 block = dim3(kBlock, 1, 1)) {
        auto key = std::make_tuple((int)k, ty, mode, colour);
        auto f = open.find(key);
        if (f != open.end()) return (*out)[f->second];
        Phase p;
        p.kind = k;
        p.ty = ty;
        p.mode = mode;
        p.colour = colour;
        p.block = block;
        out->push_back(p);
        open[key] = out->size() - 1;
        return out->back();
    }
    static BItem item(int slot) {
        BItem I{};
        I.slot = slot;
        I.gx = I.gy = I.gz = 1;
        return I;
    }
    static void add(Phase& P, BItem I, unsigned gx, unsigned gy = 1, unsigned gz = 1) {
        I.gx = gx;
        I.gy = gy;
        I.gz = gz;
        I.begin = P.blocks;
        P.blocks += (int)(gx * gy * gz);
        P.items.push_back(I);
    }
    void fin(int q, int np, int what, double tol = 0) {
        BItem I = item(q);
        I.sc = B->S[q]->sc;
        I.partial = B->S[q]->partial;
        I.i0 = np;
        I.i1 = what;
        I.tol = tol;
        Phase& P = get(K_FIN);
        I.begin = P.blocks;
        P.blocks += 1;
        P.items.push_back(I);
    }
    void vec(Kind k, int q, unsigned blocks, float* o0, const float* a0, const float* a1 = nullptr,
             float* o1 = nullptr) {
        vflip_solver* S = B->S[q];
        BItem I = item(q);
        I.n = (unsigned)S->n;
        I.o0 = o0;
        I.a0 = a0;
        I.a1 = a1;
        I.o1 = o1;
        I.sc = S->sc;
        I.partial = S->partial;
        add(get(k), I, blocks);
    }

    // One sweep of pre- (post-) smoothing number it on level l of every subdomain
    // that has a coarser level; returns through dot_np the z.r partial count of a
    // fused level-0 dot (0 when not fused).
    void sweep(int l, bool pre, int it, std::vector<int>& dot_np) {
        for (size_t q = 0; q < B->S.size(); ++q) {
            vflip_solver* S = B->S[q];
            const int nl = (int)S->L.size();
            if (l >= nl - 1) continue;
            LevelBuf& lv = S->L[l];
            const int sweeps = pre ? S->opts.pre_sweeps : S->opts.post_sweeps;
            if (it >= sweeps) continue;
            const LevelBuf* c = pre ? nullptr : &S->L[l + 1];
            const float* b = (l == 0) ? S->r : lv.b;
            BItem I = item((int)q);
            I.op = make_op_f32(lv);
            I.a0 = b;
            if (use_zsweep(S, lv)) {
                const ZGeom g = tile_geom(B, S, l);
                const int mode = (it == 0) ? (pre ? 1 : 2) : 0;
                const bool dot = !pre && l == 0 && it == sweeps - 1;
                I.a1 = x[q][l];
                I.o0 = xt[q][l];
                I.a2 = c ? x[q][l + 1] : nullptr;
                I.i0 = c ? c->ni : 0;
                I.i1 = c ? c->nj : 0;
                I.i2 = g.tiles_x;
                I.i3 = g.tiles_y;
                I.i4 = g.zchunk;
                I.partial = dot ? S->partial : nullptr;
                add(get(K_ZSWEEP, g.ty, mode), I, (unsigned)g.blocks);
                std::swap(x[q][l], xt[q][l]);
                if (dot) dot_np[q] = g.blocks;
            } else {
                I.o0 = x[q][l];
                const unsigned gb = grid_for(lv.n);
                if (it == 0) {
                    if (pre) add(get(K_FIRST), I, gb);
                    else {
                        I.a2 = x[q][l + 1];
                        I.i0 = c->ni;
                        I.i1 = c->nj;
                        add(get(K_ODD), I, gb);
                    }
                } else {
                    add(get(K_HALF, 0, 0, 1), I, gb);
                }
            }
        }
        step();
        // second half sweep of the flat levels
        for (size_t q = 0; q < B->S.size(); ++q) {
            vflip_solver* S = B->S[q];
            const int nl = (int)S->L.size();
            if (l >= nl - 1) continue;
            LevelBuf& lv = S->L[l];
            const int sweeps = pre ? S->opts.pre_sweeps : S->opts.post_sweeps;
            if (it >= sweeps || use_zsweep(S, lv)) continue;
            BItem I = item((int)q);
            I.op = make_op_f32(lv);
            I.a0 = (l == 0) ? S->r : lv.b;
            I.o0 = x[q][l];
            add(get(K_HALF, 0, 0, 0), I, grid_for(lv.n));
        }
        step();
    }

    // z = M^-1 r for every subdomain, then rho = z.r into finalize(what).
    void vcycle(int what) {
        const int ns = (int)B->S.size();
        int lmax = 0;
        for (auto* S : B->S) lmax = std::max(lmax, (int)S->L.size());
        std::vector<int> dot_np(ns, 0);
        int maxsweeps = 0;
        for (auto* S : B->S) maxsweeps = std::max({maxsweeps, S->opts.pre_sweeps, S->opts.post_sweeps});
        for (int l = 0; l < lmax - 1; ++l) {
            for (int it = 0; it < maxsweeps; ++it) sweep(l, true, it, dot_np);
            for (int q = 0; q < ns; ++q) {
                vflip_solver* S = B->S[q];
                if (l >= (int)S->L.size() - 1) continue;
                LevelBuf& lv = S->L[l];
                LevelBuf& c = S->L[l + 1];
                BItem I = item(q);
                I.op = make_op_f32(lv);
                I.a0 = (l == 0) ? S->r : lv.b;
                I.a1 = x[q][l];
                I.o0 = c.b;
                I.i0 = c.ni;
                I.i1 = c.nj;
                I.i2 = c.nk;
                add(get(K_RESTRICT, 0, 0, 0, dim3(64, 2, 2)), I, (lv.ni + 63) / 64, (lv.nj + 1) / 2, (lv.nk + 1) / 2);
            }
            step();
        }
        // every coarsest level: y = tab^T b in two passes
        for (int pass = 0; pass < 2; ++pass) {
            for (int q = 0; q < ns; ++q) {
                vflip_solver* S = B->S[q];
                const int nl = (int)S->L.size();
                LevelBuf& lv = S->L[nl - 1];
                const int nc = (int)lv.n, nch = (nc + kCoarseChunk - 1) / kCoarseChunk;
                BItem I = item(q);
                if (pass == 0) {
                    I.a0 = S->ctab;
                    I.a1 = (nl == 1) ? S->r : lv.b;
                    I.o0 = S->cpart;
                    I.i0 = nc;
                    add(get(K_CPART), I, (nc + kBlock - 1) / kBlock, nch);
                } else {
                    I.a0 = S->cpart;
                    I.o0 = x[q][nl - 1];
                    I.i0 = nc;
                    I.i1 = nch;
                    add(get(K_CRED), I, (nc + kBlock - 1) / kBlock);
                }
            }
            step();
        }
        for (int l = lmax - 2; l >= 0; --l)
            for (int it = 0; it < maxsweeps; ++it) sweep(l, false, it, dot_np);
        for (int q = 0; q < ns; ++q)
            if (!dot_np[q]) {
                vflip_solver* S = B->S[q];
                vec(K_DOT, q, S->vblocks, nullptr, x[q][0], S->r);
                dot_np[q] = S->vblocks;
            }
        step();
        for (int q = 0; q < ns; ++q) fin(q, dot_np[q], what);
        step();
    }
};

int upload(vflip_batch* B) {
    size_t total = 0;
    for (auto* list : {&B->init, &B->iter})
        for (Phase& p : *list) {
            p.off = total;
            total += p.items.size();
        }
    if (total > B->d_items_cap) {
        if (B->d_items) (void)hipFree(B->d_items);
        B->d_items = nullptr;
        B->d_items_cap = 0;
        if (hipMalloc((void**)&B->d_items, total * sizeof(BItem)) != hipSuccess) {
            set_error("hipMalloc of the batch plan failed");
            return VFLIP_ERR_NOMEM;
        }
        B->d_items_cap = total;
    }
    std::vector<BItem> all;
    all.reserve(total);
    for (auto* list : {&B->init, &B->iter})
        for (Phase& p : *list) all.insert(all.end(), p.items.begin(), p.items.end());
    VF_HIP(hipMemcpy(B->d_items, all.data(), total * sizeof(BItem), hipMemcpyHostToDevice));
    return 0;
}

void drop_graph(vflip_batch* B) {
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
}

int build_plan(vflip_batch* B) {
    drop_graph(B);
    B->init.clear();
    B->iter.clear();
    const int ns = (int)B->S.size();
    Planner P{B, nullptr, {}, {}, {}};
    P.x.resize(ns);
    P.xt.resize(ns);
    for (int q = 0; q < ns; ++q) {
        vflip_solver* S = B->S[q];
        for (size_t l = 0; l < S->L.size(); ++l) {
            P.x[q].push_back(l == 0 ? S->z : S->L[l].x);
            P.xt[q].push_back(l == 0 ? S->zt : S->L[l].xt);
        }
    }
    const auto x0 = P.x;
    // initial residual norm, x = 0, z = M^-1 r, rho, s = z
    P.out = &B->init;
    for (int q = 0; q < ns; ++q) P.vec(K_ABSMAX, q, B->S[q]->vblocks, nullptr, B->S[q]->r);
    P.step();
    for (int q = 0; q < ns; ++q) P.vec(K_FILL, q, grid_for(B->S[q]->n), B->S[q]->xw, nullptr);
    P.step();
    for (int q = 0; q < ns; ++q) P.fin(q, B->S[q]->vblocks, 4, B->S[q]->opts.tolerance);
    P.step();
    P.vcycle(0);
    for (int q = 0; q < ns; ++q) P.vec(K_COPY, q, grid_for(B->S[q]->n), B->S[q]->s, P.x[q][0]);
    P.step();
    // one PCG iteration (S/AlgebraicMultigrid.h:240-285)
    P.out = &B->iter;
    for (int q = 0; q < ns; ++q) {
        vflip_solver* S = B->S[q];
        const LevelBuf& l0 = S->L[0];
        BItem I = Planner::item(q);
        I.op = make_op_f32(l0);
        I.a0 = S->s;
        I.o0 = S->q;
        I.partial = S->partial;
        Planner::add(P.get(K_MATVEC, 0, 0, 0, dim3(64, 4, 1)), I, (l0.ni + 63) / 64, (l0.nj + 3) / 4,
                     (l0.nk + kMatvecKZ - 1) / kMatvecKZ);
    }
    P.step();
    for (int q = 0; q < ns; ++q) {
        const LevelBuf& l0 = B->S[q]->L[0];
        P.fin(q, ((l0.ni + 63) / 64) * ((l0.nj + 3) / 4) * ((l0.nk + kMatvecKZ - 1) / kMatvecKZ), 1);
    }
    P.step();
    for (int q = 0; q < ns; ++q) {
        vflip_solver* S = B->S[q];
        P.vec(K_UPXR, q, S->vblocks, S->xw, S->s, S->q, S->r);
    }
    P.step();
    for (int q = 0; q < ns; ++q) P.fin(q, B->S[q]->vblocks, 2);
    P.step();
    P.vcycle(3);
    for (int q = 0; q < ns; ++q) P.vec(K_UPS, q, B->S[q]->vblocks, B->S[q]->s, P.x[q][0]);
    P.step();
    // the sweeps ping-pong x and x-tilde; a V-cycle must leave them where it found
    // them so that every iteration (and the captured graph) uses the same buffers
    if (P.x != x0) {
        set_error("batched solve needs an even number of fused sweeps per level (pre_sweeps + post_sweeps)");
        return VFLIP_ERR_ARG;
    }
    int rc = upload(B);
    if (rc) return rc;
    B->planned = true;
    return 0;
}

int launch(vflip_batch* B, const Phase& p) {
    const BItem* it = B->d_items + p.off;
    const int n = (int)p.items.size();
    if (n == 0 || p.blocks == 0) return 0;
    hipStream_t s = B->stream;
    switch (p.kind) {
    case K_ZSWEEP:
#define VF_ZS(TY, MODE) \
    if (p.ty == TY && p.mode == MODE) kb_zsweep<TY, MODE><<<p.blocks, 256, 0, s>>>(it, n, B->d_st)
        VF_ZS(4, 0); else VF_ZS(4, 1); else VF_ZS(4, 2);
        else VF_ZS(8, 0); else VF_ZS(8, 1); else VF_ZS(8, 2);
        else VF_ZS(16, 0); else VF_ZS(16, 1); else VF_ZS(16, 2);
        else {
            set_error("no batched sweep for %d rows", p.ty);
            return VFLIP_ERR_ARG;
        }
#undef VF_ZS
        break;
    case K_FIRST: kb_first<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_HALF: kb_half<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st, p.colour); break;
    case K_ODD: kb_odd_prolong<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_RESTRICT: kb_restrict<<<p.blocks, dim3(64, 2, 2), 0, s>>>(it, n, B->d_st); break;
    case K_CPART: kb_cpart<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_CRED: kb_cred<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_MATVEC: kb_matvec<<<p.blocks, dim3(64, 4), 0, s>>>(it, n, B->d_st); break;
    case K_DOT: kb_dot<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_ABSMAX: kb_absmax<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_UPXR: kb_update_xr<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_UPS: kb_update_s<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_FILL: kb_fill<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_COPY: kb_copy<<<p.blocks, kBlock, 0, s>>>(it, n, B->d_st); break;
    case K_FIN: kb_finalize<<<p.blocks, kBlock, 0, s>>>(it, B->d_st); break;
    }
    VF_CHECK_LAUNCH();
    return 0;
}

int launch_list(vflip_batch* B, const std::vector<Phase>& list) {
    for (const Phase& p : list) {
        int rc = launch(B, p);
        if (rc) return rc;
    }
    return 0;
}

int ensure_ready(vflip_batch* B) {
    for (auto* S : B->S) {
        if (!S->setup_done || S->check_changed) {  // first use, or assembled since the last setup
            int rc = vflip_solver_setup(S);
            if (rc) return rc;
            VF_HIP(hipStreamSynchronize(S->stream));
        }
        if (!S->use_ctab) {
            set_error("batched solve needs the tabulated coarsest level (VFLIP_COARSE_TABLE unset or 1)");
            return VFLIP_ERR_ARG;
        }
        int rc = solver_ensure_ctab(S);
        if (rc) return rc;
        VF_HIP(hipStreamSynchronize(S->stream));
    }
    if (!B->planned) return build_plan(B);
    return 0;
}

int capture(vflip_batch* B) {
    if (B->exec) return 0;
    VF_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < B->graph_iters && !rc; ++k) rc = launch_list(B, B->iter);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return VFLIP_ERR_HIP;
    }
    B->graph = g;
    VF_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    return 0;
}

}  // namespace

extern "C" {

int vflip_batch_create(int n, vflip_solver* const* solvers, vflip_batch** out) {
    VF_ARG(out && n > 0 && solvers, "batch arguments");
    *out = nullptr;
    for (int q = 0; q < n; ++q) {
        VF_ARG(solvers[q], "solver");
        VF_ARG(solvers[q]->opts.precision == VFLIP_PREC_FP32, "batched solves are FP32 (plain) solvers");
        VF_ARG(solvers[q]->opts.pre_sweeps > 0 && solvers[q]->opts.post_sweeps > 0, "batched solves need sweeps");
    }
    vflip_batch* B = new vflip_batch();
    B->S.assign(solvers, solvers + n);
    int rc = 0;
    // non-blocking: the batch orders itself against the solvers' streams explicitly (it
    // synchronises them on entry and its own stream on exit)
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&B->e0) != hipSuccess ||
        hipEventCreate(&B->e1) != hipSuccess) {
        set_error("hipStreamCreate/hipEventCreate failed");
        rc = VFLIP_ERR_HIP;
    }
    if (!rc && (hipMalloc((void**)&B->d_st, n * sizeof(BState)) != hipSuccess ||
                hipMalloc((void**)&B->d_maxit, n * sizeof(int)) != hipSuccess ||
                hipHostMalloc((void**)&B->h_st, n * sizeof(BState), hipHostMallocDefault) != hipSuccess)) {
        set_error("batch state allocation failed");
        rc = VFLIP_ERR_NOMEM;
    }
    if (!rc) {
        std::vector<int> m(n);
        for (int q = 0; q < n; ++q) m[q] = solvers[q]->opts.max_iterations;
        if (hipMemcpy(B->d_maxit, m.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("hipMemcpy failed");
            rc = VFLIP_ERR_HIP;
        }
    }
    if (rc) {
        vflip_batch_destroy(B);
        return rc;
    }
    *out = B;
    return 0;
}

int vflip_batch_destroy(vflip_batch* B) {
    if (!B) return 0;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    drop_graph(B);
    if (B->d_items) (void)hipFree(B->d_items);
    if (B->d_st) (void)hipFree(B->d_st);
    if (B->d_maxit) (void)hipFree(B->d_maxit);
    if (B->h_st) (void)hipHostFree(B->h_st);
    if (B->e0) (void)hipEventDestroy(B->e0);
    if (B->e1) (void)hipEventDestroy(B->e1);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return 0;
}

int vflip_batch_set_graph_iterations(vflip_batch* B, int iterations) {
    VF_ARG(B && iterations > 0 && iterations <= 64, "graph iterations (1..64)");
    if (iterations != B->graph_iters) {
        drop_graph(B);
        B->graph_iters = iterations;
    }
    return 0;
}

}  // extern "C"

namespace {
// finish: outputs and stats as vflip_solver_solve reports them (FP64 true residual per
// subdomain); without it only x and the iteration counts (its, may be null), for the
// correction solves of vflip_batch_solve_refined.
int batch_solve_impl(vflip_batch* B, const float* const* rhs, float* const* x, vflip_solve_stats* stats,
                     int use_graph, double* ms, bool finish, int* its) {
    const int ns = (int)B->S.size();
    for (int q = 0; q < ns; ++q) VF_ARG(rhs[q] && x[q], "rhs/x of a subdomain");
    // the solvers' own streams may hold earlier work on these buffers (assembly)
    for (auto* S : B->S) VF_HIP(hipStreamSynchronize(S->stream));
    int rc = ensure_ready(B);
    if (rc) return rc;
    if (use_graph && (rc = capture(B))) return rc;
    VF_HIP(hipEventRecord(B->e0, B->stream));
    for (int q = 0; q < ns; ++q)
        VF_HIP(hipMemcpyAsync(B->S[q]->r, rhs[q], B->S[q]->n * sizeof(float), hipMemcpyDeviceToDevice, B->stream));
    kb_reset<<<(ns + 63) / 64, 64, 0, B->stream>>>(B->d_st, B->d_maxit, ns);
    VF_CHECK_LAUNCH();
    if ((rc = launch_list(B, B->init))) return rc;
    int maxit = 0;
    for (auto* S : B->S) maxit = std::max(maxit, S->opts.max_iterations);
    for (int done_its = 0; done_its < maxit;) {
        if (use_graph) {
            VF_HIP(hipGraphLaunch(B->exec, B->stream));
            done_its += B->graph_iters;
        } else {
            const int k = std::min(B->graph_iters, maxit - done_its);
            for (int i = 0; i < k && !rc; ++i) rc = launch_list(B, B->iter);
            if (rc) return rc;
            done_its += k;
        }
        VF_HIP(hipMemcpyAsync(B->h_st, B->d_st, ns * sizeof(BState), hipMemcpyDeviceToHost, B->stream));
        VF_HIP(hipStreamSynchronize(B->stream));
        bool all = true;
        for (int q = 0; q < ns; ++q) all = all && B->h_st[q].done;
        if (all) break;
    }
    VF_HIP(hipEventRecord(B->e1, B->stream));
    VF_HIP(hipMemcpyAsync(B->h_st, B->d_st, ns * sizeof(BState), hipMemcpyDeviceToHost, B->stream));
    VF_HIP(hipStreamSynchronize(B->stream));
    if (ms) {
        float t = 0.f;
        VF_HIP(hipEventElapsedTime(&t, B->e0, B->e1));
        *ms = t;
    }
    int worst = VFLIP_OK;
    for (int q = 0; q < ns; ++q) {
        const BState& st = B->h_st[q];
        const int status = st.done ? st.status : VFLIP_NOT_CONVERGED;
        if (finish) {
            if ((rc = solver_finish(B->S[q], rhs[q], x[q], st.iters, status, stats ? &stats[q] : nullptr))) return rc;
        } else {
            VF_HIP(hipMemcpyAsync(x[q], B->S[q]->xw, B->S[q]->n * sizeof(float), hipMemcpyDeviceToDevice, B->stream));
            if (its) its[q] = st.iters;
        }
        if (status < 0) worst = status;
        else if (worst == VFLIP_OK && status != VFLIP_OK) worst = status;
    }
    return worst;
}
}  // namespace

extern "C" {

int vflip_batch_solve(vflip_batch* B, const float* const* rhs, float* const* x, vflip_solve_stats* stats,
                      int use_graph, double* ms) {
    VF_ARG(B && rhs && x, "batch/rhs/x");
    return batch_solve_impl(B, rhs, x, stats, use_graph, ms, true, nullptr);
}

int vflip_batch_solve_refined(vflip_batch* B, const float* const* rhs, float* const* x, double* const* x64,
                              int assembled, double tol, int max_rounds, vflip_solve_stats* stats, int use_graph,
                              double* ms) {
    VF_ARG(B && rhs && x, "batch/rhs/x");
    VF_ARG(tol > 0 && max_rounds >= 0, "tolerance / rounds");
    const int ns = (int)B->S.size();
    for (int q = 0; q < ns; ++q) VF_ARG(rhs[q] && x[q], "rhs/x of a subdomain");
    for (auto* S : B->S) VF_HIP(hipStreamSynchronize(S->stream));
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = B->stream;
    int rc;
    for (auto* S : B->S)
        if ((rc = solver_refine_begin(S, s, assembled != 0))) return rc;
    std::vector<int> its(ns, 0);
    std::vector<double> bnorm(ns, 0.0), tr(ns, 0.0);
    std::vector<char> done(ns, 0);
    std::vector<const float*> in_rhs(ns);
    std::vector<float*> in_x(ns);
    std::vector<int> in_its(ns, 0);
    for (int round = 0;; ++round) {
        // FP64 true residual of every unfinished subdomain, one read-back for all
        for (int q = 0; q < ns; ++q)
            if (!done[q]) {
                vflip_solver* S = B->S[q];
                if ((rc = solver_refine_residual(S, s, rhs[q], assembled != 0))) return rc;
                VF_HIP(hipMemcpyAsync(S->h_sc, S->sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
            }
        VF_HIP(hipStreamSynchronize(s));
        bool all = true;
        for (int q = 0; q < ns; ++q) {
            if (done[q]) continue;
            tr[q] = B->S[q]->h_sc->true_res;
            if (round == 0) bnorm[q] = tr[q];
            if (bnorm[q] == 0 || tr[q] <= tol * bnorm[q]) done[q] = 1;
            all = all && done[q];
        }
        if (all || round == max_rounds) break;
        // one batched FP32 solve of the corrections; finished subdomains get a zero
        // right-hand side, which their solve takes as converged at once
        for (int q = 0; q < ns; ++q) {
            vflip_solver* S = B->S[q];
            if (done[q]) k_fill<<<grid_for(S->n), kBlock, 0, s>>>((unsigned)S->n, S->ref_r, 0.f);
            in_rhs[q] = S->ref_r;
            in_x[q] = S->ref_e;
        }
        VF_CHECK_LAUNCH();
        VF_HIP(hipStreamSynchronize(s));
        rc = batch_solve_impl(B, in_rhs.data(), in_x.data(), nullptr, use_graph, nullptr, false, in_its.data());
        if (rc < 0) return rc;
        for (int q = 0; q < ns; ++q)
            if (!done[q]) {
                its[q] += in_its[q];
                if ((rc = solver_refine_add(B->S[q], s))) return rc;
            }
    }
    int worst = VFLIP_OK;
    for (int q = 0; q < ns; ++q) {
        if ((rc = solver_refine_end(B->S[q], s, x[q], x64 ? x64[q] : nullptr))) return rc;
        if (!done[q]) worst = VFLIP_NOT_CONVERGED;
    }
    VF_HIP(hipStreamSynchronize(s));
    const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms) *ms = t;
    if (stats)
        for (int q = 0; q < ns; ++q) solver_refine_stats(its[q], done[q] != 0, tr[q], bnorm[q], t, &stats[q]);
    return worst;
}

int vflip_batch_autotune(vflip_batch* B, int* out_ty, int* out_zc, int max_levels) {
    VF_ARG(B, "batch");
    int rc = ensure_ready(B);
    if (rc) return rc;
    int lmax = 0;
    for (auto* S : B->S) lmax = std::max(lmax, (int)S->L.size());
    B->tune.assign(lmax, Tune{});
    for (int l = 0; l < lmax; ++l) {
        // subdomains whose level l takes the fused sweep
        std::vector<int> idx;
        int nj_min = 1 << 30, nk_min = 1 << 30;
        for (int q = 0; q < (int)B->S.size(); ++q) {
            vflip_solver* S = B->S[q];
            if (l >= (int)S->L.size() - 1 || !use_zsweep(S, S->L[l])) continue;
            idx.push_back(q);
            nj_min = std::min(nj_min, S->L[l].nj);
            nk_min = std::min(nk_min, S->L[l].nk);
        }
        if (idx.empty()) continue;
        float best = 1e30f;
        Tune pick{};
        for (int ty : {4, 8, 16}) {
            if (ty > nj_min) continue;
            for (int zc : {4, 8, 16, 32, 64}) {
                if (zc > nk_min && zc != 4) continue;
                B->tune[l] = Tune{ty, zc};
                std::vector<Phase> one;
                Phase p;
                p.kind = K_ZSWEEP;
                p.ty = ty;
                p.mode = 0;
                bool fits = true;
                for (int q : idx) {
                    vflip_solver* S = B->S[q];
                    LevelBuf& lv = S->L[l];
                    const ZGeom g = tile_geom(B, S, l);
                    if (l == 0 && g.blocks > S->np_partial) fits = false;  // the fused dot's partials
                    BItem I = Planner::item(q);
                    I.op = make_op_f32(lv);
                    I.a0 = (l == 0) ? S->r : lv.b;
                    I.a1 = (l == 0) ? S->z : lv.x;
                    I.o0 = (l == 0) ? S->zt : lv.xt;
                    I.i2 = g.tiles_x;
                    I.i3 = g.tiles_y;
                    I.i4 = g.zchunk;
                    Planner::add(p, I, (unsigned)g.blocks);
                }
                if (!fits) continue;
                one.push_back(p);
                // time it on its own item array
                B->init.swap(one);
                B->iter.clear();
                if ((rc = upload(B))) return rc;
                VF_HIP(hipMemsetAsync(B->d_st, 0, B->S.size() * sizeof(BState), B->stream));
                for (int w = 0; w < 2; ++w)
                    if ((rc = launch(B, B->init[0]))) return rc;
                VF_HIP(hipEventRecord(B->e0, B->stream));
                const int reps = 10;
                for (int w = 0; w < reps; ++w)
                    if ((rc = launch(B, B->init[0]))) return rc;
                VF_HIP(hipEventRecord(B->e1, B->stream));
                VF_HIP(hipEventSynchronize(B->e1));
                float t = 0.f;
                VF_HIP(hipEventElapsedTime(&t, B->e0, B->e1));
                B->init.swap(one);
                if (t < best) {
                    best = t;
                    pick = Tune{ty, zc};
                }
            }
        }
        B->tune[l] = pick;
        if (l < max_levels) {
            if (out_ty) out_ty[l] = pick.ty;
            if (out_zc) out_zc[l] = pick.zc;
        }
    }
    B->planned = false;
    return build_plan(B);
}

}  // extern "C"
