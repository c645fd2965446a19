// Pressure-projection kernels around the solver: cut-cell face weights, inflow
// boundary, matrix-free assembly and the gradient update.
//
// Face weights and cell indices must match the reference bit for bit, so this
// file is compiled without FMA contraction and with IEEE fp32 division.
#pragma clang fp contract(off)

#include <vector>

#include "common.h"

namespace vflip {

// fraction_inside for a segment, S/levelset_util.cpp:4-13
__device__ inline float frac2(float l, float r) {
    if (l < 0 && r < 0) return 1.f;
    if (l < 0 && r >= 0) return l / (l - r);
    if (l >= 0 && r < 0) return r / (r - l);
    return 0.f;
}

// fraction_inside for a square, S/levelset_util.cpp:23-103. Corners are walked
// bl, br, tr, tl and rotated until the case's anchor corner comes first.
__device__ inline float frac4(float bl, float br, float tl, float tr) {
    int inside = (bl <= 0 ? 1 : 0) + (tl <= 0 ? 1 : 0) + (br <= 0 ? 1 : 0) + (tr <= 0 ? 1 : 0);
    float c0 = bl, c1 = br, c2 = tr, c3 = tl;
    auto rot = [&]() {
        float t = c0;
        c0 = c1; c1 = c2; c2 = c3; c3 = t;
    };
    if (inside == 4) return 1.f;
    if (inside == 3) {
        for (int q = 0; q < 4 && c0 < 0; ++q) rot();
        float s0 = 1 - frac2(c0, c3);
        float s1 = 1 - frac2(c0, c1);
        return 1 - 0.5f * s0 * s1;
    }
    if (inside == 2) {
        for (int q = 0; q < 4 && (c0 >= 0 || !(c1 < 0 || c2 < 0)); ++q) rot();
        if (c1 < 0) {
            float sl = frac2(c0, c3);
            float sr = frac2(c1, c2);
            return 0.5f * (sl + sr);
        }
        float mid = 0.25f * (c0 + c1 + c2 + c3);
        if (mid < 0) {
            float area = 0;
            float s1 = 1 - frac2(c0, c3);
            float s3 = 1 - frac2(c2, c3);
            area += 0.5f * s1 * s3;
            float s2 = 1 - frac2(c2, c1);
            float s0 = 1 - frac2(c0, c1);
            area += 0.5f * s0 * s2;
            return 1 - area;
        }
        float area = 0;
        float s0 = frac2(c0, c1);
        float s1 = frac2(c0, c3);
        area += 0.5f * s0 * s1;
        float s2 = frac2(c2, c1);
        float s3 = frac2(c2, c3);
        area += 0.5f * s2 * s3;
        return area;
    }
    if (inside == 1) {
        for (int q = 0; q < 4 && c0 >= 0; ++q) rot();
        float s0 = frac2(c0, c3);
        float s1 = frac2(c0, c1);
        return 0.5f * s0 * s1;
    }
    return 0.f;
}

__device__ inline float clamp01(float a) { return a < 0.f ? 0.f : (a > 1.f ? 1.f : a); }

// axis 0: u faces (ni+1,nj,nk), corners (i,j,k),(i,j+1,k),(i,j,k+1),(i,j+1,k+1)
// axis 1: v faces (ni,nj+1,nk), corners (i,j,k),(i,j,k+1),(i+1,j,k),(i+1,j,k+1)
// axis 2: w faces (ni,nj,nk+1), corners (i,j,k),(i,j+1,k),(i+1,j,k),(i+1,j+1,k)
// S/Smoke_solver3D.cpp:882-941
__global__ __launch_bounds__(kBlock) void k_face_weights(int axis, int ni, int nj, int nk, const float* __restrict__ phi,
                                                         float* __restrict__ out) {
    const int fx = ni + (axis == 0), fy = nj + (axis == 1), fz = nk + (axis == 2);
    const size_t nf = (size_t)fx * fy * fz;
    size_t f = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (f >= nf) return;
    int i = (int)(f % fx);
    size_t t = f / fx;
    int j = (int)(t % fy), k = (int)(t / fy);
    const int px = ni + 1, py = nj + 1;
    auto P = [&](int a, int b, int c) { return phi[lin3(a, b, c, px, py)]; };
    float fr;
    if (axis == 0) fr = frac4(P(i, j, k), P(i, j + 1, k), P(i, j, k + 1), P(i, j + 1, k + 1));
    else if (axis == 1) fr = frac4(P(i, j, k), P(i, j, k + 1), P(i + 1, j, k), P(i + 1, j, k + 1));
    else fr = frac4(P(i, j, k), P(i, j + 1, k), P(i + 1, j, k), P(i + 1, j + 1, k));
    out[f] = clamp01(1 - fr);
}

// S/Smoke_solver3D.cpp:964-980
__global__ __launch_bounds__(kBlock) void k_inflow(int ni, int nj, int nk, const int8_t* __restrict__ bdesc, float uin,
                                                   float* __restrict__ u, float* __restrict__ v, float* __restrict__ w) {
    const size_t n = (size_t)ni * nj * nk;
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n || bdesc[c] != 1) return;
    int i = (int)(c % ni);
    size_t t = c / ni;
    int j = (int)(t % nj), k = (int)(t / nj);
    u[lin3(i, j, k, ni + 1, nj)] = uin;
    u[lin3(i + 1, j, k, ni + 1, nj)] = uin;
    v[lin3(i, j, k, ni, nj + 1)] = 0.f;
    v[lin3(i, j + 1, k, ni, nj + 1)] = 0.f;
    w[lin3(i, j, k, ni, nj)] = 0.f;
    w[lin3(i, j, k + 1, ni, nj)] = 0.f;
}

// Matrix-free assembly over interior cells 1..n-2 (ring rows empty).
// S/Smoke_solver3D.cpp:986-1111; sub variant S/sub_domain.cpp:612-729.
// Writes diag/ax/ay/az/leak (FP32, from FP64 accumulation of the FP32 terms)
// and rhs (FP64 accumulation rounded to FP32). With a liquid level set (lphi !=
// null; liquid where lphi < 0) an air cell is a Dirichlet p = 0 cell like the
// ring (empty row) and a liquid cell couples only to liquid neighbours: the
// first-order free-surface boundary (no reference counterpart: the reference's
// smoke assembly is the all-liquid case).
__global__ __launch_bounds__(kBlock) void k_assemble(int ni, int nj, int nk, float h, float dt, int sub,
                                                     const float* __restrict__ lphi,
                                                     const float* __restrict__ u, const float* __restrict__ v,
                                                     const float* __restrict__ w, const float* __restrict__ uw,
                                                     const float* __restrict__ vw, const float* __restrict__ ww,
                                                     const float* __restrict__ us, const float* __restrict__ vs,
                                                     const float* __restrict__ ws, float* __restrict__ D,
                                                     float* __restrict__ AX, float* __restrict__ AY,
                                                     float* __restrict__ AZ, float* __restrict__ LK,
                                                     float* __restrict__ rhs, int* __restrict__ changed,
                                                     double* __restrict__ D64, double* __restrict__ R64) {
    const size_t n = (size_t)ni * nj * nk;
    size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    int i = (int)(c % ni);
    size_t t = c / ni;
    int j = (int)(t % nj), k = (int)(t / nj);
    auto wet = [&](size_t q) { return !lphi || lphi[q] < 0.f; };
    if (!(i >= 1 && i < ni - 1 && j >= 1 && j < nj - 1 && k >= 1 && k < nk - 1) || !wet(c)) {
        // with change detection the coefficients are stored only where they differ
        // (an unchanged operator costs no coefficient writes)
        const bool diff = __float_as_uint(D[c]) | __float_as_uint(AX[c]) | __float_as_uint(AY[c]) |
                          __float_as_uint(AZ[c]) | __float_as_uint(LK[c]);
        if (changed && diff) *changed = 1;  // every writer stores the same value
        if (!changed || diff) { D[c] = 0.f; AX[c] = 0.f; AY[c] = 0.f; AZ[c] = 0.f; LK[c] = 0.f; }
        rhs[c] = 0.f;
        if (D64) { D64[c] = 0.0; R64[c] = 0.0; }
        return;
    }
    const float h2 = h * h;
    const size_t fxp = lin3(i + 1, j, k, ni + 1, nj), fxm = lin3(i, j, k, ni + 1, nj);
    const size_t fyp = lin3(i, j + 1, k, ni, nj + 1), fym = lin3(i, j, k, ni, nj + 1);
    const size_t fzp = lin3(i, j, k + 1, ni, nj), fzm = lin3(i, j, k, ni, nj);
    auto term = [&](float wgt) { return sub ? wgt * dt / h2 : wgt / h2; };
    const float txp = term(uw[fxp]), txm = term(uw[fxm]);
    const float typ = term(vw[fyp]), tym = term(vw[fym]);
    const float tzp = term(ww[fzp]), tzm = term(ww[fzm]);
    double d = 0, r = 0, lk = 0;
    d += txp; r -= (double)(uw[fxp] * u[fxp] / h);
    d += txm; r += (double)(uw[fxm] * u[fxm] / h);
    d += typ; r -= (double)(vw[fyp] * v[fyp] / h);
    d += tym; r += (double)(vw[fym] * v[fym] / h);
    d += tzp; r -= (double)(ww[fzp] * w[fzp] / h);
    d += tzm; r += (double)(ww[fzm] * w[fzm] / h);
    // terms whose neighbour is a ring (or air) cell carry no off-diagonal (Dirichlet leak)
    const size_t sl = (size_t)ni * nj;
    const bool oxp = i < ni - 2 && wet(c + 1), oxm = i > 1 && wet(c - 1);
    const bool oyp = j < nj - 2 && wet(c + ni), oym = j > 1 && wet(c - ni);
    const bool ozp = k < nk - 2 && wet(c + sl), ozm = k > 1 && wet(c - sl);
    if (!oxp) lk += txp;
    if (!oxm) lk += txm;
    if (!oyp) lk += typ;
    if (!oym) lk += tym;
    if (!ozp) lk += tzp;
    if (!ozm) lk += tzm;
    if (d > 0) {
        float s = (uw[fxp] - 1.0f) * us[fxp];
        s = s - (uw[fxm] - 1.0f) * us[fxm];
        s = s + (vw[fyp] - 1.0f) * vs[fyp];
        s = s - (vw[fym] - 1.0f) * vs[fym];
        s = s + (ww[fzp] - 1.0f) * ws[fzp];
        s = s - (ww[fzm] - 1.0f) * ws[fzm];
        if (sub) r += (double)((1 / h) * s);
        else r += 1.0 / (double)h * (double)s;
    }
    const float nd = (float)d, nx = oxp ? -txp : 0.f, ny = oyp ? -typ : 0.f;
    const float nz = ozp ? -tzp : 0.f, nl = (float)lk;
    // flag an operator change against the previous assembly (bitwise compare)
    const bool diff = !changed ||
                      (__float_as_uint(D[c]) != __float_as_uint(nd) || __float_as_uint(AX[c]) != __float_as_uint(nx) ||
                       __float_as_uint(AY[c]) != __float_as_uint(ny) || __float_as_uint(AZ[c]) != __float_as_uint(nz) ||
                       __float_as_uint(LK[c]) != __float_as_uint(nl));
    if (changed && diff) *changed = 1;  // every writer stores the same value
    if (D64) {
        D64[c] = d;  // the reference's FP64 diagonal and right-hand side
        R64[c] = r;
    }
    if (diff) {  // bit-equal coefficients are already stored
        D[c] = nd;
        AX[c] = nx;
        AY[c] = ny;
        AZ[c] = nz;
        LK[c] = nl;
    }
    rhs[c] = (float)r;
}

// u -= grad p on the three MAC components in one pass (S/Smoke_solver3D.cpp:1126-1200;
// SUB: S/sub_domain.cpp:745-819, dt/h scaling and faces 1..n-1 along the axis). A thread
// takes lattice point (i, j, k) of the (ni+1) x (nj+1) x (nk+1) box and the u, v and w
// faces of that index that exist, so each pressure value is read once for the three
// faces on its low side. Per face: along the component's axis the faces lo..hi
// (COARSE 2..n-2, SUB 1..n-1), across it 1..n-2; where the face weight is > 0 the
// velocity drops (float)(p[c] - p[c - e]) / h (times dt for SUB), else it takes the
// solid velocity. P = double reproduces the reference exactly (its pressure vector is
// double); P = float is the plain FP32 path.
template <typename P>
__device__ inline void grad_face(int axis, int i, int j, int k, int ni, int nj, int nk, int sub, float h, float dt,
                                 const P* __restrict__ p, const float* __restrict__ wgt,
                                 const float* __restrict__ sol, float* __restrict__ vel) {
    const int fx = ni + (axis == 0), fy = nj + (axis == 1);
    const int a[3] = {i, j, k}, nn[3] = {ni, nj, nk};
    const int lo = sub ? 1 : 2;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q == axis) {
            const int hi = sub ? nn[q] - 1 : nn[q] - 2;
            if (a[q] < lo || a[q] > hi) return;
        } else if (a[q] < 1 || a[q] > nn[q] - 2) {
            return;
        }
    }
    const size_t f = ((size_t)k * fy + j) * fx + i;
    if (wgt[f] > 0) {
        const size_t c = lin3(i, j, k, ni, nj);
        const size_t cm = (axis == 0) ? c - 1 : (axis == 1) ? c - ni : c - (size_t)ni * nj;
        const float g = (float)(p[c] - p[cm]);
        vel[f] -= sub ? dt * g / h : g / h;
    } else {
        vel[f] = sol[f];
    }
}

template <typename P>
__global__ __launch_bounds__(256) void k_grad3(int ni, int nj, int nk, int sub, float h, float dt,
                                               const P* __restrict__ p, const float* __restrict__ uw,
                                               const float* __restrict__ vw, const float* __restrict__ ww,
                                               const float* __restrict__ us, const float* __restrict__ vs,
                                               const float* __restrict__ ws, float* __restrict__ u,
                                               float* __restrict__ v, float* __restrict__ w) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i > ni || j > nj || k > nk) return;
    grad_face<P>(0, i, j, k, ni, nj, nk, sub, h, dt, p, uw, us, u);
    grad_face<P>(1, i, j, k, ni, nj, nk, sub, h, dt, p, vw, vs, v);
    grad_face<P>(2, i, j, k, ni, nj, nk, sub, h, dt, p, ww, ws, w);
}

}  // namespace vflip

using namespace vflip;

extern "C" {

int vflip_face_weights(int ni, int nj, int nk, const float* phi, float* uw, float* vw, float* ww, void* stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && phi && uw && vw && ww, "face_weights args");
    hipStream_t s = (hipStream_t)stream;
    size_t nu = (size_t)(ni + 1) * nj * nk, nv = (size_t)ni * (nj + 1) * nk, nw = (size_t)ni * nj * (nk + 1);
    k_face_weights<<<grid_for(nu), kBlock, 0, s>>>(0, ni, nj, nk, phi, uw);
    k_face_weights<<<grid_for(nv), kBlock, 0, s>>>(1, ni, nj, nk, phi, vw);
    k_face_weights<<<grid_for(nw), kBlock, 0, s>>>(2, ni, nj, nk, phi, ww);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

int vflip_inflow_bc(int ni, int nj, int nk, const int8_t* bdesc, float u_in, float* u, float* v, float* w,
                    void* stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && bdesc && u && v && w, "inflow args");
    hipStream_t s = (hipStream_t)stream;
    k_inflow<<<grid_for((size_t)ni * nj * nk), kBlock, 0, s>>>(ni, nj, nk, bdesc, u_in, u, v, w);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"

namespace {
// World position of point (i,j,k) of a grid with Buffer3D offsets o, in the
// reference's float arithmetic: Vec3f(i,j,k)*h - Vec3f(ox,oy,oz)*h
// (S/Smoke_solver3D.cpp:806,864). The file's contract(off) keeps the host
// compiler from fusing the product into the subtraction.
void grid_point(int i, int j, int k, float h, const float o[3], float* q) {
    const float a = o[0] * h, b = o[1] * h, c = o[2] * h;
    q[0] = (float)i * h - a;
    q[1] = (float)j * h - b;
    q[2] = (float)k * h - c;
}

// Samples one component of the callback at every point of an nx x ny x nz grid
// (comp < 0: the scalar callback) and uploads the values.
int sample_grid(int nx, int ny, int nz, float h, const float o[3], vflip_phi_fn phi, vflip_vel_fn vel, int comp,
                void* user, float* dst, hipStream_t s) {
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> host(n);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                float q[3], r[3] = {0.f, 0.f, 0.f};
                grid_point(i, j, k, h, o, q);
                float val;
                if (comp < 0) {
                    val = phi(q, user);
                } else {
                    vel(q, r, user);
                    val = r[comp];
                }
                host[lin3(i, j, k, nx, ny)] = val;
            }
    VF_HIP(hipMemcpyAsync(dst, host.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    VF_HIP(hipStreamSynchronize(s));
    return 0;
}
}  // namespace

extern "C" {

int vflip_set_boundary_phi(int ni, int nj, int nk, float h, vflip_phi_fn phi, void* user, float* nodal_phi,
                           void* stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && h > 0 && phi && nodal_phi, "set_boundary_phi args");
    const float o[3] = {0.5f, 0.5f, 0.5f};
    return sample_grid(ni + 1, nj + 1, nk + 1, h, o, phi, nullptr, -1, user, nodal_phi, (hipStream_t)stream);
}

int vflip_set_boundary_vel(int ni, int nj, int nk, float h, vflip_vel_fn vel, void* user, float* us, float* vs,
                           float* ws, void* stream) {
    VF_ARG(ni > 0 && nj > 0 && nk > 0 && h > 0 && vel && us && vs && ws, "set_boundary_vel args");
    hipStream_t s = (hipStream_t)stream;
    const float ou[3] = {0.5f, 0.f, 0.f}, ov[3] = {0.f, 0.5f, 0.f}, ow[3] = {0.f, 0.f, 0.5f};
    int rc = sample_grid(ni + 1, nj, nk, h, ou, nullptr, vel, 0, user, us, s);
    if (!rc) rc = sample_grid(ni, nj + 1, nk, h, ov, nullptr, vel, 1, user, vs, s);
    if (!rc) rc = sample_grid(ni, nj, nk + 1, h, ow, nullptr, vel, 2, user, ws, s);
    return rc;
}

int vflip_grad_apply(int ni, int nj, int nk, int variant, float h, float dt, const float* p, const float* uw,
                     const float* vw, const float* ww, const float* us, const float* vs, const float* ws, float* u,
                     float* v, float* w, void* stream) {
    VF_ARG(ni > 2 && nj > 2 && nk > 2, "grid dims");
    VF_ARG(p && uw && vw && ww && us && vs && ws && u && v && w, "grad arrays");
    VF_ARG(variant == VFLIP_VARIANT_COARSE || variant == VFLIP_VARIANT_SUB, "variant");
    hipStream_t s = (hipStream_t)stream;
    int sub = variant == VFLIP_VARIANT_SUB;
    k_grad3<float><<<dim3(ni / 64 + 1, nj / 4 + 1, nk + 1), dim3(64, 4), 0, s>>>(ni, nj, nk, sub, h, dt, p, uw, vw, ww,
                                                                              us, vs, ws, u, v, w);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_grad_apply_f64(int ni, int nj, int nk, int variant, float h, float dt, const double* p, const float* uw,
                         const float* vw, const float* ww, const float* us, const float* vs, const float* ws,
                         float* u, float* v, float* w, void* stream) {
    VF_ARG(ni > 2 && nj > 2 && nk > 2, "grid dims");
    VF_ARG(p && uw && vw && ww && us && vs && ws && u && v && w, "grad arrays");
    VF_ARG(variant == VFLIP_VARIANT_COARSE || variant == VFLIP_VARIANT_SUB, "variant");
    hipStream_t s = (hipStream_t)stream;
    int sub = variant == VFLIP_VARIANT_SUB;
    k_grad3<double><<<dim3(ni / 64 + 1, nj / 4 + 1, nk + 1), dim3(64, 4), 0, s>>>(ni, nj, nk, sub, h, dt, p, uw, vw,
                                                                               ww, us, vs, ws, u, v, w);
    VF_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"

// Assembly and the whole projection need the solver's level-0 storage.
struct vflip_solver;
namespace vflip {
int solver_level0(vflip_solver* s, int* ni, int* nj, int* nk, float** d, float** ax, float** ay, float** az,
                  float** leak, hipStream_t* stream);
void solver_mark_matrix_changed(vflip_solver* s);
int* solver_begin_assembly(vflip_solver* s);
int solver_f64(vflip_solver* s, double** d64, double** b64, double** x64);
void solver_mark_f64_assembled(vflip_solver* s);
bool solver_is_mixed(const vflip_solver* s);
int solver_set_warm_start(vflip_solver* s, int on);
int solver_solve_assembled(vflip_solver* s, const float* rhs32, float* x32, vflip_solve_stats* st);
int solver_history_guess(vflip_solver* S, const float* rhs32, double** guess, int* used);
int solver_guess_residual(vflip_solver* S, const float* rhs32, const double* g, float** r);
float* solver_scratch32(vflip_solver* S);
int solver_add_correction(vflip_solver* S, double* g, const float* corr);
int solver_solve_history(vflip_solver* S, const float* rhs32, double* g, float* x32, vflip_solve_stats* st, bool warm);
int solver_guess_residual_restricted(vflip_solver* S, const float* rhs32, const double* g, float* crhs,
                                     float* csolid, float* cliquid);
int unet_warm_inputs(vflip_unet* N, int cx, int cy, int cz, float** crhs, float** csolid, float** cliquid);
int unet_warm_predict_add(vflip_unet* N, int ni, int nj, int nk, int factor, const float* diag, double* g);
}  // namespace vflip

extern "C" {

static int assemble_impl(vflip_solver* S, int variant, float h, float dt, const float* u, const float* v,
                         const float* w, const float* uw, const float* vw, const float* ww, const float* us,
                         const float* vs, const float* ws, const float* liquid_phi, float* rhs) {
    VF_ARG(S, "solver");
    VF_ARG(u && v && w && uw && vw && ww && us && vs && ws && rhs, "assembly arrays");
    VF_ARG(variant == VFLIP_VARIANT_COARSE || variant == VFLIP_VARIANT_SUB, "variant");
    int ni, nj, nk;
    float *d, *ax, *ay, *az, *lk;
    hipStream_t s;
    int rc = solver_level0(S, &ni, &nj, &nk, &d, &ax, &ay, &az, &lk, &s);
    if (rc) return rc;
    double *d64, *b64, *x64;
    rc = solver_f64(S, &d64, &b64, &x64);
    if (rc) return rc;
    int* changed = solver_begin_assembly(S);
    k_assemble<<<grid_for((size_t)ni * nj * nk), kBlock, 0, s>>>(ni, nj, nk, h, dt, variant == VFLIP_VARIANT_SUB,
                                                                 liquid_phi, u, v, w, uw, vw, ww, us, vs, ws, d, ax,
                                                                 ay, az, lk, rhs, changed, d64, b64);
    VF_CHECK_LAUNCH();
    solver_mark_f64_assembled(S);
    return 0;
}

int vflip_assemble(vflip_solver* S, int variant, float h, float dt, const float* u, const float* v, const float* w,
                   const float* uw, const float* vw, const float* ww, const float* us, const float* vs,
                   const float* ws, float* rhs) {
    return assemble_impl(S, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, nullptr, rhs);
}

int vflip_assemble_liquid(vflip_solver* S, float h, float dt, const float* u, const float* v, const float* w,
                          const float* uw, const float* vw, const float* ww, const float* us, const float* vs,
                          const float* ws, const float* liquid_phi, float* rhs) {
    VF_ARG(liquid_phi, "liquid level set");
    return assemble_impl(S, VFLIP_VARIANT_COARSE, h, dt, u, v, w, uw, vw, ww, us, vs, ws, liquid_phi, rhs);
}

namespace {
int apply_pressure(vflip_solver* S, int variant, float h, float dt, float* u, float* v, float* w, const float* uw,
                   const float* vw, const float* ww, const float* us, const float* vs, const float* ws,
                   const float* pressure);

int project_impl(vflip_solver* S, vflip_unet* net, int factor, int variant, float h, float dt, float* u, float* v,
                 float* w, const float* uw, const float* vw, const float* ww, const float* us, const float* vs,
                 const float* ws, float* rhs_scratch, float* pressure, vflip_solve_stats* stats) {
    VF_ARG(S && pressure && rhs_scratch, "solver/pressure/rhs");
    int rc = vflip_assemble(S, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs_scratch);
    if (rc) return rc;
    rc = vflip_solver_setup(S);
    if (rc) return rc;
    int status;
    if (net) {
        // the network's pressure guess seeds the PCG, which still runs to the tolerance
        int ni, nj, nk;
        float *d, *ax, *ay, *az, *lk;
        hipStream_t s;
        rc = solver_level0(S, &ni, &nj, &nk, &d, &ax, &ay, &az, &lk, &s);
        if (rc) return rc;
        if ((rc = vflip_unet_set_stream(net, s))) return rc;
        if ((rc = vflip_unet_warm_start(net, ni, nj, nk, factor, rhs_scratch, d, pressure))) return rc;
        const int old = solver_set_warm_start(S, 1);
        status = solver_solve_assembled(S, rhs_scratch, pressure, stats);
        solver_set_warm_start(S, old);
    } else {
        status = solver_solve_assembled(S, rhs_scratch, pressure, stats);
    }
    if (status < 0 && status != VFLIP_ERR_BREAKDOWN) return status;
    rc = apply_pressure(S, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, pressure);
    if (rc) return rc;
    return status;
}

// Pressure-history projection (vflip_project_history): guess from the solver's FP64
// history, optionally corrected by the network from the divergence the guess leaves.
int project_history_impl(vflip_solver* S, vflip_unet* net, int factor, int variant, float h, float dt, float* u,
                         float* v, float* w, const float* uw, const float* vw, const float* ww, const float* us,
                         const float* vs, const float* ws, float* rhs_scratch, float* pressure,
                         vflip_solve_stats* stats) {
    VF_ARG(S && pressure && rhs_scratch, "solver/pressure/rhs");
    int rc = vflip_assemble(S, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs_scratch);
    if (rc) return rc;
    rc = vflip_solver_setup(S);
    if (rc) return rc;
    double* g = nullptr;
    int used = 0;
    if ((rc = solver_history_guess(S, rhs_scratch, &g, &used))) return rc;
    bool warm = used > 0;
    if (net) {
        // the correction from the network, on the residual the guess leaves: restricted
        // by 4 straight from the guess (no full-resolution residual), prolonged back
        // and added into the FP64 guess
        int ni, nj, nk;
        float *d, *ax, *ay, *az, *lk;
        hipStream_t s;
        if ((rc = solver_level0(S, &ni, &nj, &nk, &d, &ax, &ay, &az, &lk, &s))) return rc;
        if ((rc = vflip_unet_set_stream(net, s))) return rc;
        if (factor == 4 && ni % 4 == 0 && nj % 4 == 0 && nk % 4 == 0) {
            float *crhs, *csolid, *cliquid;
            if ((rc = unet_warm_inputs(net, ni / 4, nj / 4, nk / 4, &crhs, &csolid, &cliquid))) return rc;
            if ((rc = solver_guess_residual_restricted(S, rhs_scratch, g, crhs, csolid, cliquid))) return rc;
            if ((rc = unet_warm_predict_add(net, ni, nj, nk, 4, d, g))) return rc;
        } else {
            float* r = nullptr;
            if ((rc = solver_guess_residual(S, rhs_scratch, g, &r))) return rc;
            float* corr = solver_scratch32(S);
            if ((rc = vflip_unet_warm_start(net, ni, nj, nk, factor, r, d, corr))) return rc;
            if ((rc = solver_add_correction(S, g, corr))) return rc;
        }
        warm = true;
    }
    const int status = solver_solve_history(S, rhs_scratch, g, pressure, stats, warm);
    if (status < 0 && status != VFLIP_ERR_BREAKDOWN) return status;
    rc = apply_pressure(S, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, pressure);
    if (rc) return rc;
    return status;
}
}  // namespace

namespace {
// u -= grad p after a solve: the FP64 pressure in mixed mode, as the reference
// (double pressure vector), else the FP32 one.
int apply_pressure(vflip_solver* S, int variant, float h, float dt, float* u, float* v, float* w, const float* uw,
                   const float* vw, const float* ww, const float* us, const float* vs, const float* ws,
                   const float* pressure) {
    int rc;
    int ni, nj, nk;
    float *d, *ax, *ay, *az, *lk;
    hipStream_t s;
    rc = solver_level0(S, &ni, &nj, &nk, &d, &ax, &ay, &az, &lk, &s);
    if (rc) return rc;
    if (solver_is_mixed(S)) {
        // gradient of the FP64 pressure, as the reference (double pressure vector)
        double *d64, *b64, *x64;
        rc = solver_f64(S, &d64, &b64, &x64);
        if (rc) return rc;
        rc = vflip_grad_apply_f64(ni, nj, nk, variant, h, dt, x64, uw, vw, ww, us, vs, ws, u, v, w, s);
    } else {
        rc = vflip_grad_apply(ni, nj, nk, variant, h, dt, pressure, uw, vw, ww, us, vs, ws, u, v, w, s);
    }
    return rc;
}
}  // namespace

int vflip_project(vflip_solver* S, int variant, float h, float dt, float* u, float* v, float* w, const float* uw,
                  const float* vw, const float* ww, const float* us, const float* vs, const float* ws,
                  float* rhs_scratch, float* pressure, vflip_solve_stats* stats) {
    return project_impl(S, nullptr, 1, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs_scratch, pressure, stats);
}

int vflip_project_learned(vflip_solver* S, vflip_unet* net, int factor, int variant, float h, float dt, float* u,
                          float* v, float* w, const float* uw, const float* vw, const float* ww, const float* us,
                          const float* vs, const float* ws, float* rhs_scratch, float* pressure,
                          vflip_solve_stats* stats) {
    VF_ARG(net, "network");
    return project_impl(S, net, factor, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs_scratch, pressure, stats);
}

int vflip_project_history(vflip_solver* S, vflip_unet* net, int factor, int variant, float h, float dt, float* u,
                          float* v, float* w, const float* uw, const float* vw, const float* ww, const float* us,
                          const float* vs, const float* ws, float* rhs_scratch, float* pressure,
                          vflip_solve_stats* stats) {
    return project_history_impl(S, net, factor, variant, h, dt, u, v, w, uw, vw, ww, us, vs, ws, rhs_scratch,
                                pressure, stats);
}

}  // extern "C"
