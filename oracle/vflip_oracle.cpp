// vflip oracle: CPU FP64 restatement of the reference vFLIP per-step hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP product path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and only to check or to time the reference algorithm on the host. The
// product library (vflip_amd/libvflip.so) never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root, S/ = Smoke_Solver/).
//
// Parity status of each function (how it is pinned; see tests/):
//   fraction_inside2/4, face_weights .... oracle/_ref (reference levelset_util.cpp
//                                         compiled as-is) + closed-form areas
//   sample_linear/cubic/cube_lerp ....... oracle/_ref (reference fluid_buffer3D.h
//                                         compiled as-is) + affine/cubic exactness
//   randhash ............................ oracle/_ref (reference util.h)
//   assemble ............................ closed form (uniform Laplacian) + scipy
//                                         divergence/Laplacian operators
//   coarsen ............................. scipy 0.5*R*A*P with the reference's R/P
//   matvec/rbgs/vcycle/amgpcg ........... scipy spsolve, RB-GS colour-order brute
//                                         force, PCG invariants
//   grad_apply .......................... discrete divergence after projection, coarse
//                                         and subdomain ranges (all interior cells in
//                                         the sub form; faces outside the ranges untouched)
//   assemble, solid-velocity term ....... closed form u = u_solid affine: rhs = -div u_solid
//                                         at every cell with diag > 0, any face weights
//   inflow_bc ........................... closed form (explicit face lists)
//   grid_points ......................... oracle/_ref (the reference's Vec3f arithmetic)
//                                         + closed form (i - o) h
//   hash (binning) ...................... brute-force membership
//   p2g ................................. brute force over all particles; constant
//                                         velocity closed form
//   g2p ................................. affine reproduction
//   advect_rk2_linear/cubic_clamp ....... affine translation / RK2 exactness
//   coarse/fine operators ............... constant-field round trip, flux closed form
//   liquid_phi .......................... brute-force distance over all particles
//   reorder ............................. brute-force restatement of the per-cell
//                                         rules on tiny grids, affine velocity and
//                                         plane solid closed forms, idempotence;
//                                         the reseed RNG replaces rand() (no vector
//                                         can pin rand() across platforms)
//   nested coupling: get_velocity blend . closed forms (coarse value outside is_in2,
//                                         subdomain value deep inside, the 4h ramp)
//   advect_particles_nested ............. translation equals the plain coarse
//                                         advection; plane push-out lands on the plane
//   halo_select ......................... brute-force membership and order
//   seed_ghost .......................... plane-solid closed forms (cells, bands,
//                                         solid velocity behind, blend in front)
//   tracers: advect ..................... translation, decay closed form, plain vs
//                                         nested-without-subdomain identity
//   tracers: emitters, frame ............ inside-sphere/count, phi band with a
//                                         plane, frame record closed form
//   pse_diffuse ......................... float64 brute force over the 27 cells,
//                                         uniform velocity unchanged, far particles
//                                         unchanged, kernel value closed form
//
// Floating point: compiled with -ffp-contract=off so float expressions round
// exactly as written (the reference's MSVC /fp:precise build does the same).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline size_t lin(int i, int j, int k, int nx, int ny) {
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
}

// S/util.h:173 clamp
template <class T> inline T clampv(T a, T lo, T hi) {
    if (a < lo) return lo;
    if (a > hi) return hi;
    return a;
}

// Truncating float->int conversion of Vec3i(Vec3f) (S/vec.h:37-39) with the
// x86-64 result for values outside int range and NaN (the "integer indefinite"
// INT_MIN of cvttss2si), written out so it does not rest on undefined behaviour.
inline int trunc_i32(float t) {
    if (!(t > -2147483648.0f && t < 2147483648.0f)) return INT32_MIN;
    return (int)t;
}

// S/levelset_util.cpp:15-20 cycle_array
inline void rot4(float* a) {
    float t = a[0];
    a[0] = a[1]; a[1] = a[2]; a[2] = a[3]; a[3] = t;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Level-set cut fractions. S/levelset_util.cpp:4-13 (segment) and 23-103 (square)
// ---------------------------------------------------------------------------
float or_fraction_inside2(float l, float r) {
    if (l < 0 && r < 0) return 1;
    if (l < 0 && r >= 0) return l / (l - r);
    if (l >= 0 && r < 0) return r / (r - l);
    return 0;
}

// The reference rotates with unbounded while loops; with corners exactly 0 some
// of them never terminate (e.g. two zeros and two positives). The rotations are
// bounded at one full turn here, which is identical wherever the reference ends.
float or_fraction_inside4(float bl, float br, float tl, float tr) {
    int inside = (bl <= 0 ? 1 : 0) + (tl <= 0 ? 1 : 0) + (br <= 0 ? 1 : 0) + (tr <= 0 ? 1 : 0);
    // corner order walked around the square: bl, br, tr, tl
    float c[4] = {bl, br, tr, tl};
    if (inside == 4) return 1;
    if (inside == 3) {
        for (int q = 0; q < 4 && c[0] < 0; ++q) rot4(c);
        float s0 = 1 - or_fraction_inside2(c[0], c[3]);
        float s1 = 1 - or_fraction_inside2(c[0], c[1]);
        return 1 - 0.5f * s0 * s1;
    }
    if (inside == 2) {
        for (int q = 0; q < 4 && (c[0] >= 0 || !(c[1] < 0 || c[2] < 0)); ++q) rot4(c);
        if (c[1] < 0) {
            float sl = or_fraction_inside2(c[0], c[3]);
            float sr = or_fraction_inside2(c[1], c[2]);
            return 0.5f * (sl + sr);
        }
        float mid = 0.25f * (c[0] + c[1] + c[2] + c[3]);
        if (mid < 0) {
            float area = 0;
            float s1 = 1 - or_fraction_inside2(c[0], c[3]);
            float s3 = 1 - or_fraction_inside2(c[2], c[3]);
            area += 0.5f * s1 * s3;
            float s2 = 1 - or_fraction_inside2(c[2], c[1]);
            float s0 = 1 - or_fraction_inside2(c[0], c[1]);
            area += 0.5f * s0 * s2;
            return 1 - area;
        }
        float area = 0;
        float s0 = or_fraction_inside2(c[0], c[1]);
        float s1 = or_fraction_inside2(c[0], c[3]);
        area += 0.5f * s0 * s1;
        float s2 = or_fraction_inside2(c[2], c[1]);
        float s3 = or_fraction_inside2(c[2], c[3]);
        area += 0.5f * s2 * s3;
        return area;
    }
    if (inside == 1) {
        for (int q = 0; q < 4 && c[0] >= 0; ++q) rot4(c);
        float s0 = or_fraction_inside2(c[0], c[3]);
        float s1 = or_fraction_inside2(c[0], c[1]);
        return 0.5f * s0 * s1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Cut-cell face weights from the nodal solid SDF.
// S/Smoke_solver3D.cpp:874-944 (identical in S/sub_domain.cpp:491-554).
// phi: nodal (ni+1)x(nj+1)x(nk+1); uw: (ni+1)xnjxnk; vw: nix(nj+1)xnk; ww: nixnjx(nk+1).
// All arrays linear, x fastest.
// ---------------------------------------------------------------------------
void or_face_weights(int ni, int nj, int nk, const float* phi, float* uw, float* vw, float* ww) {
    const int px = ni + 1, py = nj + 1;
    auto P = [&](int i, int j, int k) { return phi[lin(i, j, k, px, py)]; };
#pragma omp parallel for
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i <= ni; ++i) {
                float f = 1 - or_fraction_inside4(P(i, j, k), P(i, j + 1, k), P(i, j, k + 1), P(i, j + 1, k + 1));
                uw[lin(i, j, k, ni + 1, nj)] = clampv(f, 0.0f, 1.0f);
            }
#pragma omp parallel for
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j <= nj; ++j)
            for (int i = 0; i < ni; ++i) {
                float f = 1 - or_fraction_inside4(P(i, j, k), P(i, j, k + 1), P(i + 1, j, k), P(i + 1, j, k + 1));
                vw[lin(i, j, k, ni, nj + 1)] = clampv(f, 0.0f, 1.0f);
            }
#pragma omp parallel for
    for (int k = 0; k <= nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                float f = 1 - or_fraction_inside4(P(i, j, k), P(i, j + 1, k), P(i + 1, j, k), P(i + 1, j + 1, k));
                ww[lin(i, j, k, ni, nj)] = clampv(f, 0.0f, 1.0f);
            }
}

// ---------------------------------------------------------------------------
// Pressure system assembly over interior cells 1..n-2 (outer ring = empty rows,
// Dirichlet p=0). S/Smoke_solver3D.cpp:983-1111 (coarse, sub=0) and
// S/sub_domain.cpp:612-729 (sub=1: terms carry dt, solid term uses float 1/h).
// Output: symmetric 7-point matrix as diag + couplings to +x/+y/+z neighbours
// (ax[c] = A(c, c+1) = A(c+1, c), etc.; zero where the reference stores no entry),
// plus the FP64 right-hand side. All nixnjxnk linear.
// ---------------------------------------------------------------------------
// liquid (NULL for smoke): per-cell 1 = liquid, 0 = air. An air cell is a
// Dirichlet p = 0 cell like the domain ring (empty row), and a liquid cell
// couples only to liquid neighbours: the first-order free-surface boundary of
// the standard MAC projection (the reference has no liquid solver; its smoke
// assembly is the all-liquid case).
void or_assemble(int ni, int nj, int nk, float h, float dt, int sub,
                 const float* u, const float* v, const float* w,
                 const float* uw, const float* vw, const float* ww,
                 const float* us, const float* vs, const float* ws, const float* liquid,
                 double* diag, double* ax, double* ay, double* az, double* rhs) {
    const size_t n = (size_t)ni * nj * nk;
    std::memset(diag, 0, n * sizeof(double));
    std::memset(ax, 0, n * sizeof(double));
    std::memset(ay, 0, n * sizeof(double));
    std::memset(az, 0, n * sizeof(double));
    std::memset(rhs, 0, n * sizeof(double));
    auto U = [&](const float* a, int i, int j, int k) { return a[lin(i, j, k, ni + 1, nj)]; };
    auto V = [&](const float* a, int i, int j, int k) { return a[lin(i, j, k, ni, nj + 1)]; };
    auto W = [&](const float* a, int i, int j, int k) { return a[lin(i, j, k, ni, nj)]; };
    const float h2 = h * h;  // sqr(_hx), S/util.h:26
#pragma omp parallel for
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i) {
                size_t c = lin(i, j, k, ni, nj);
                if (liquid && !(liquid[c] > 0.f)) continue;
                auto wet = [&](size_t q) { return !liquid || liquid[q] > 0.f; };
                double d = 0, r = 0;
                float t;
                // +x face
                t = sub ? U(uw, i + 1, j, k) * dt / h2 : U(uw, i + 1, j, k) / h2;
                d += t;
                if (i < ni - 2 && wet(c + 1)) ax[c] = -(double)t;
                r -= (double)(U(uw, i + 1, j, k) * U(u, i + 1, j, k) / h);
                // -x face
                t = sub ? U(uw, i, j, k) * dt / h2 : U(uw, i, j, k) / h2;
                d += t;
                r += (double)(U(uw, i, j, k) * U(u, i, j, k) / h);
                // +y face
                t = sub ? V(vw, i, j + 1, k) * dt / h2 : V(vw, i, j + 1, k) / h2;
                d += t;
                if (j < nj - 2 && wet(c + (size_t)ni)) ay[c] = -(double)t;
                r -= (double)(V(vw, i, j + 1, k) * V(v, i, j + 1, k) / h);
                // -y face
                t = sub ? V(vw, i, j, k) * dt / h2 : V(vw, i, j, k) / h2;
                d += t;
                r += (double)(V(vw, i, j, k) * V(v, i, j, k) / h);
                // +z face
                t = sub ? W(ww, i, j, k + 1) * dt / h2 : W(ww, i, j, k + 1) / h2;
                d += t;
                if (k < nk - 2 && wet(c + (size_t)ni * nj)) az[c] = -(double)t;
                r -= (double)(W(ww, i, j, k + 1) * W(w, i, j, k + 1) / h);
                // -z face
                t = sub ? W(ww, i, j, k) * dt / h2 : W(ww, i, j, k) / h2;
                d += t;
                r += (double)(W(ww, i, j, k) * W(w, i, j, k) / h);
                if (d > 0) {
                    // solid-velocity term, evaluated left to right in float
                    float s = (U(uw, i + 1, j, k) - 1.0f) * U(us, i + 1, j, k);
                    s = s - (U(uw, i, j, k) - 1.0f) * U(us, i, j, k);
                    s = s + (V(vw, i, j + 1, k) - 1.0f) * V(vs, i, j + 1, k);
                    s = s - (V(vw, i, j, k) - 1.0f) * V(vs, i, j, k);
                    s = s + (W(ww, i, j, k + 1) - 1.0f) * W(ws, i, j, k + 1);
                    s = s - (W(ww, i, j, k) - 1.0f) * W(ws, i, j, k);
                    if (sub)
                        r += (double)((1 / h) * s);  // S/sub_domain.cpp:721: int/float
                    else
                        r += 1.0 / (double)h * (double)s;  // S/Smoke_solver3D.cpp:1101: double
                }
                diag[c] = d;
                rhs[c] = r;
            }
}

// ---------------------------------------------------------------------------
// AMG hierarchy shapes: ceil-halving while unknowns > 16^3.
// S/GeometricLevelGen.h:36-58.
// ---------------------------------------------------------------------------
int or_level_dims(int ni, int nj, int nk, int* dims, int max_levels) {
    int L = 1;
    dims[0] = ni; dims[1] = nj; dims[2] = nk;
    unsigned int unknowns = (unsigned)ni * nj * nk;
    int a = ni, b = nj, c = nk;
    while (unknowns > 16u * 16u * 16u && L < max_levels) {
        a = (int)std::ceil((float)a / 2.0);
        b = (int)std::ceil((float)b / 2.0);
        c = (int)std::ceil((float)c / 2.0);
        dims[3 * L] = a; dims[3 * L + 1] = b; dims[3 * L + 2] = c;
        unknowns = (unsigned)a * b * c;
        ++L;
    }
    return L;
}

// ---------------------------------------------------------------------------
// Galerkin coarsening A_c = 0.5 * R * A * P with R = 1/8 per child (average)
// and P = injection, S/GeometricLevelGen.cpp:8-76 and S/GeometricLevelGen.h:60-68.
// Piecewise-constant aggregates only touch across faces, so A_c is again a
// symmetric 7-point operator; its entries in closed form are
//   A_c(I,I)   = 1/16 * sum_{i,k in I} A(i,k)
//   A_c(I,I+e) = 1/16 * sum_{i in I, k in I+e} A(i,k)
// which is what the sparse triple product evaluates (up to FP64 summation order).
// ---------------------------------------------------------------------------
void or_coarsen(int ni, int nj, int nk, const double* d, const double* ax, const double* ay,
                const double* az, double* dc, double* axc, double* ayc, double* azc) {
    const int ci = (int)std::ceil((float)ni / 2.0), cj = (int)std::ceil((float)nj / 2.0),
              ck = (int)std::ceil((float)nk / 2.0);
    const double s = 0.5 * 0.125;  // 0.5 * R weight * P weight (1.0)
#pragma omp parallel for
    for (int K = 0; K < ck; ++K)
        for (int J = 0; J < cj; ++J)
            for (int I = 0; I < ci; ++I) {
                double sd = 0, sx = 0, sy = 0, sz = 0;
                for (int kk = 0; kk < 2; ++kk)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int ii = 0; ii < 2; ++ii) {
                            int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                            if (i >= ni || j >= nj || k >= nk) continue;
                            size_t c = lin(i, j, k, ni, nj);
                            sd += d[c];
                            // couplings from child i to a neighbour k
                            if (i + 1 < ni) { if (ii == 0) sd += 2 * ax[c]; else sx += ax[c]; }
                            if (j + 1 < nj) { if (jj == 0) sd += 2 * ay[c]; else sy += ay[c]; }
                            if (k + 1 < nk) { if (kk == 0) sd += 2 * az[c]; else sz += az[c]; }
                        }
                size_t C = lin(I, J, K, ci, cj);
                dc[C] = s * sd;
                axc[C] = (I + 1 < ci) ? s * sx : 0.0;
                ayc[C] = (J + 1 < cj) ? s * sy : 0.0;
                azc[C] = (K + 1 < ck) ? s * sz : 0.0;
            }
}

// 7-point operator application in the reference CSR column order
// (-z, -y, -x, diag, +x, +y, +z). S/sparse_matrix.h:341-366.
static inline double row_dot(int i, int j, int k, int ni, int nj, int nk, const double* d,
                             const double* ax, const double* ay, const double* az, const double* x,
                             bool with_diag) {
    size_t c = lin(i, j, k, ni, nj);
    double s = 0;
    if (k > 0) s += az[c - (size_t)ni * nj] * x[c - (size_t)ni * nj];
    if (j > 0) s += ay[c - ni] * x[c - ni];
    if (i > 0) s += ax[c - 1] * x[c - 1];
    if (with_diag) s += d[c] * x[c];
    if (i + 1 < ni) s += ax[c] * x[c + 1];
    if (j + 1 < nj) s += ay[c] * x[c + ni];
    if (k + 1 < nk) s += az[c] * x[c + (size_t)ni * nj];
    return s;
}

void or_matvec(int ni, int nj, int nk, const double* d, const double* ax, const double* ay,
               const double* az, const double* x, double* y) {
#pragma omp parallel for
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i)
                y[lin(i, j, k, ni, nj)] = row_dot(i, j, k, ni, nj, nk, d, ax, ay, az, x, true);
}

// Red-black Gauss-Seidel, odd colour (i+j+k odd) first. S/AlgebraicMultigrid.h:22-103.
void or_rbgs(int ni, int nj, int nk, const double* d, const double* ax, const double* ay,
             const double* az, const double* b, double* x, int sweeps) {
    for (int it = 0; it < sweeps; ++it) {
        for (int colour = 1; colour >= 0; --colour) {
#pragma omp parallel for
            for (int k = 0; k < nk; ++k)
                for (int j = 0; j < nj; ++j)
                    for (int i = 0; i < ni; ++i) {
                        if (((i + j + k) & 1) != colour) continue;
                        size_t c = lin(i, j, k, ni, nj);
                        double s = row_dot(i, j, k, ni, nj, nk, d, ax, ay, az, x, false);
                        x[c] = (d[c] != 0) ? (b[c] - s) / d[c] : 0.0;
                    }
        }
    }
}

// b_c = R (b - A x), residual formed entry by entry as multiply_and_subtract does.
// S/AlgebraicMultigrid.h:105-117, S/sparse_matrix.h:459-480.
void or_restrict_residual(int ni, int nj, int nk, const double* d, const double* ax,
                          const double* ay, const double* az, const double* x, const double* b,
                          double* bc) {
    const int ci = (int)std::ceil((float)ni / 2.0), cj = (int)std::ceil((float)nj / 2.0),
              ck = (int)std::ceil((float)nk / 2.0);
    std::vector<double> r((size_t)ni * nj * nk);
#pragma omp parallel for
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                size_t c = lin(i, j, k, ni, nj);
                double rr = b[c];
                const size_t sl = (size_t)ni * nj;
                if (k > 0) rr -= az[c - sl] * x[c - sl];
                if (j > 0) rr -= ay[c - ni] * x[c - ni];
                if (i > 0) rr -= ax[c - 1] * x[c - 1];
                rr -= d[c] * x[c];
                if (i + 1 < ni) rr -= ax[c] * x[c + 1];
                if (j + 1 < nj) rr -= ay[c] * x[c + ni];
                if (k + 1 < nk) rr -= az[c] * x[c + sl];
                r[c] = rr;
            }
#pragma omp parallel for
    for (int K = 0; K < ck; ++K)
        for (int J = 0; J < cj; ++J)
            for (int I = 0; I < ci; ++I) {
                double s = 0;
                for (int kk = 0; kk < 2; ++kk)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int ii = 0; ii < 2; ++ii) {
                            int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                            if (i < ni && j < nj && k < nk) s += 0.125 * r[lin(i, j, k, ni, nj)];
                        }
                bc[lin(I, J, K, ci, cj)] = s;
            }
}

// x += P x_c (injection). S/AlgebraicMultigrid.h:118-129.
void or_prolong_add(int ni, int nj, int nk, const double* xc, double* x) {
    const int ci = (int)std::ceil((float)ni / 2.0), cj = (int)std::ceil((float)nj / 2.0);
#pragma omp parallel for
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) x[lin(i, j, k, ni, nj)] += xc[lin(i / 2, j / 2, k / 2, ci, cj)];
}

}  // extern "C"

namespace {

struct Level {
    int n[3];
    std::vector<double> d, ax, ay, az, x, b;
};

struct Hierarchy {
    std::vector<Level> L;
};

void build_hierarchy(Hierarchy& H, int ni, int nj, int nk, const double* d, const double* ax,
                     const double* ay, const double* az) {
    int dims[3 * 32];
    int nl = or_level_dims(ni, nj, nk, dims, 32);
    H.L.resize(nl);
    for (int l = 0; l < nl; ++l) {
        Level& lv = H.L[l];
        lv.n[0] = dims[3 * l]; lv.n[1] = dims[3 * l + 1]; lv.n[2] = dims[3 * l + 2];
        size_t n = (size_t)lv.n[0] * lv.n[1] * lv.n[2];
        lv.d.resize(n); lv.ax.resize(n); lv.ay.resize(n); lv.az.resize(n);
        lv.x.assign(n, 0.0); lv.b.assign(n, 0.0);
        if (l == 0) {
            std::copy(d, d + n, lv.d.begin());
            std::copy(ax, ax + n, lv.ax.begin());
            std::copy(ay, ay + n, lv.ay.begin());
            std::copy(az, az + n, lv.az.begin());
        } else {
            Level& f = H.L[l - 1];
            or_coarsen(f.n[0], f.n[1], f.n[2], f.d.data(), f.ax.data(), f.ay.data(), f.az.data(),
                       lv.d.data(), lv.ax.data(), lv.ay.data(), lv.az.data());
        }
    }
}

// One V-cycle applied to r with zero initial guess: z = M^-1 r.
// S/AlgebraicMultigrid.h:130-188 (4 pre / 4 post sweeps, 200 on the coarsest level).
void vcycle(Hierarchy& H, const double* r, double* z, int pre, int post, int coarse) {
    int nl = (int)H.L.size();
    for (int l = 0; l < nl; ++l) std::fill(H.L[l].x.begin(), H.L[l].x.end(), 0.0);
    std::copy(r, r + H.L[0].b.size(), H.L[0].b.begin());
    for (int l = 0; l < nl - 1; ++l) {
        Level& a = H.L[l];
        or_rbgs(a.n[0], a.n[1], a.n[2], a.d.data(), a.ax.data(), a.ay.data(), a.az.data(), a.b.data(),
                a.x.data(), pre);
        or_restrict_residual(a.n[0], a.n[1], a.n[2], a.d.data(), a.ax.data(), a.ay.data(), a.az.data(),
                             a.x.data(), a.b.data(), H.L[l + 1].b.data());
    }
    {
        Level& a = H.L[nl - 1];
        or_rbgs(a.n[0], a.n[1], a.n[2], a.d.data(), a.ax.data(), a.ay.data(), a.az.data(), a.b.data(),
                a.x.data(), coarse);
    }
    for (int l = nl - 2; l >= 0; --l) {
        Level& a = H.L[l];
        or_prolong_add(a.n[0], a.n[1], a.n[2], H.L[l + 1].x.data(), a.x.data());
        or_rbgs(a.n[0], a.n[1], a.n[2], a.d.data(), a.ax.data(), a.ay.data(), a.az.data(), a.b.data(),
                a.x.data(), post);
    }
    std::copy(H.L[0].x.begin(), H.L[0].x.end(), z);
}

// serial dot and abs-max, S/blas_wrapper.h:13-43
double sdot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}
double absmax(const std::vector<double>& a) {
    double m = 0;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::fabs(a[i]) > m) m = std::fabs(a[i]);
    return m;
}

}  // namespace

extern "C" {

// One V-cycle preconditioner application (for kernel-level parity checks).
void or_vcycle(int ni, int nj, int nk, const double* d, const double* ax, const double* ay,
               const double* az, const double* r, double* z, int pre, int post, int coarse) {
    Hierarchy H;
    build_hierarchy(H, ni, nj, nk, d, ax, ay, az);
    vcycle(H, r, z, pre, post, coarse);
}

// Aggregation-AMG preconditioned CG with the reference's infinity-norm stop
// |r|_inf <= tol * |b|_inf on the recursively updated residual.
// S/AlgebraicMultigrid.h:189-301. Returns 1 converged, 0 not converged.
// history (optional, length maxit+1): |r|_inf after each iteration (history[0] = |b|_inf).
int or_amgpcg(int ni, int nj, int nk, const double* d, const double* ax, const double* ay,
              const double* az, const double* rhs, double* result, double tol, int maxit,
              double* residual_out, int* iterations_out, int pre, int post, int coarse,
              double* history) {
    const size_t n = (size_t)ni * nj * nk;
    Hierarchy H;
    build_hierarchy(H, ni, nj, nk, d, ax, ay, az);
    std::vector<double> r(rhs, rhs + n), z(n), s(n), x(n, 0.0);
    double res = absmax(r);
    *residual_out = res;
    if (history) history[0] = res;
    if (res == 0) {
        *iterations_out = 0;
        std::fill(result, result + n, 0.0);
        return 1;
    }
    double tol_abs = tol * res;
    vcycle(H, r.data(), z.data(), pre, post, coarse);
    double rho = sdot(z, r);
    if (rho == 0 || rho != rho) {
        *iterations_out = 0;
        std::fill(result, result + n, 0.0);
        return 0;
    }
    s = z;
    int it;
    for (it = 0; it < maxit; ++it) {
        or_matvec(ni, nj, nk, d, ax, ay, az, s.data(), z.data());
        double alpha = rho / sdot(s, z);
#pragma omp parallel for
        for (long long q = 0; q < (long long)n; ++q) {
            x[q] += alpha * s[q];
            r[q] += -alpha * z[q];
        }
        res = absmax(r);
        if (history) history[it + 1] = res;
        if (res <= tol_abs) {
            *residual_out = res;
            *iterations_out = it + 1;
            std::copy(x.begin(), x.end(), result);
            return 1;
        }
        vcycle(H, r.data(), z.data(), pre, post, coarse);
        double rho_new = sdot(z, r);
        double beta = rho_new / rho;
#pragma omp parallel for
        for (long long q = 0; q < (long long)n; ++q) s[q] = z[q] + beta * s[q];
        rho = rho_new;
    }
    *residual_out = res;
    *iterations_out = it;
    std::copy(x.begin(), x.end(), result);
    return 0;
}

// ---------------------------------------------------------------------------
// Inflow boundary written before the assembly, S/Smoke_solver3D.cpp:964-980:
// every cell with bdesc == 1 sets its two x faces to u_in and its y and z faces
// to 0. Faces shared with a neighbour get the same value from either cell.
// ---------------------------------------------------------------------------
void or_inflow_bc(int ni, int nj, int nk, const signed char* bdesc, float u_in, float* u, float* v, float* w) {
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                if (bdesc[lin(i, j, k, ni, nj)] != 1) continue;
                u[lin(i, j, k, ni + 1, nj)] = u_in;
                u[lin(i + 1, j, k, ni + 1, nj)] = u_in;
                v[lin(i, j, k, ni, nj + 1)] = 0;
                v[lin(i, j + 1, k, ni, nj + 1)] = 0;
                w[lin(i, j, k, ni, nj)] = 0;
                w[lin(i, j, k + 1, ni, nj)] = 0;
            }
}

// ---------------------------------------------------------------------------
// World positions at which set_boundary_phi / set_boundary_vel sample their
// callbacks: Vec3f(i,j,k)*h - Vec3f(ox,oy,oz)*h in float, per component
// (S/Smoke_solver3D.cpp:806,822,838 with the face grids' offsets (0.5,0,0),
// (0,0.5,0), (0,0,0.5); 864 with (0.5,0.5,0.5) for the nodal SDF).
// out: 3 floats per point of the nx x ny x nz grid, x fastest.
// ---------------------------------------------------------------------------
void or_grid_points(int nx, int ny, int nz, float h, float ox, float oy, float oz, float* out) {
    const float o[3] = {ox * h, oy * h, oz * h};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                float* q = out + 3 * lin(i, j, k, nx, ny);
                q[0] = (float)i * h - o[0];
                q[1] = (float)j * h - o[1];
                q[2] = (float)k * h - o[2];
            }
}

// ---------------------------------------------------------------------------
// Pressure-gradient update. S/Smoke_solver3D.cpp:1126-1200 (sub=0: faces
// 2..n-2 along the face axis, no dt) and S/sub_domain.cpp:745-819 (sub=1: faces
// 1..n-1, multiplied by dt). Faces with weight 0 take the solid velocity.
// ---------------------------------------------------------------------------
void or_grad_apply(int ni, int nj, int nk, float h, float dt, int sub, const double* p,
                   const float* uw, const float* vw, const float* ww, const float* us,
                   const float* vs, const float* ws, float* u, float* v, float* w) {
    const int lo = sub ? 1 : 2;
    const size_t sl = (size_t)ni * nj;
#pragma omp parallel for
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = lo; i <= (sub ? ni - 1 : ni - 2); ++i) {
                size_t f = lin(i, j, k, ni + 1, nj), c = lin(i, j, k, ni, nj);
                if (uw[f] > 0) {
                    float g = (float)(p[c] - p[c - 1]);
                    u[f] -= sub ? dt * g / h : g / h;
                } else {
                    u[f] = us[f];
                }
            }
#pragma omp parallel for
    for (int k = 1; k < nk - 1; ++k)
        for (int j = lo; j <= (sub ? nj - 1 : nj - 2); ++j)
            for (int i = 1; i < ni - 1; ++i) {
                size_t f = lin(i, j, k, ni, nj + 1), c = lin(i, j, k, ni, nj);
                if (vw[f] > 0) {
                    float g = (float)(p[c] - p[c - ni]);
                    v[f] -= sub ? dt * g / h : g / h;
                } else {
                    v[f] = vs[f];
                }
            }
#pragma omp parallel for
    for (int k = lo; k <= (sub ? nk - 1 : nk - 2); ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i) {
                size_t f = lin(i, j, k, ni, nj), c = lin(i, j, k, ni, nj);
                if (ww[f] > 0) {
                    float g = (float)(p[c] - p[c - sl]);
                    w[f] -= sub ? dt * g / h : g / h;
                } else {
                    w[f] = ws[f];
                }
            }
}

// ---------------------------------------------------------------------------
// Samplers on a linear (x-fastest) nx x ny x nz grid with grid offset (ox,oy,oz)
// and spacing h; index clamping as Buffer3D::at. Interpolation in FP64.
// S/fluid_buffer3D.h:205-248 (at, sample_linear), 259-326 (sample_cubic),
// 329-356 (sample_cube_lerp), 22-29 (cubic weights).
// ---------------------------------------------------------------------------
static inline float at_c(const float* f, int nx, int ny, int nz, int i, int j, int k) {
    int ti = std::min(std::max(i, 0), nx - 1), tj = std::min(std::max(j, 0), ny - 1),
        tk = std::min(std::max(k, 0), nz - 1);
    return f[lin(ti, tj, tk, nx, ny)];
}
static inline double lerpd(double a, double b, double c) { return (1.0 - c) * a + c * b; }

float or_sample_linear(const float* f, int nx, int ny, int nz, double h, double ox, double oy,
                       double oz, double wx, double wy, double wz) {
    double gx = wx / h + ox, gy = wy / h + oy, gz = wz / h + oz;
    int i = (int)std::floor(gx), j = (int)std::floor(gy), k = (int)std::floor(gz);
    double cx = gx - (double)i, cy = gy - (double)j, cz = gz - (double)k;
    double v1 = lerpd(lerpd(at_c(f, nx, ny, nz, i, j, k), at_c(f, nx, ny, nz, i + 1, j, k), cx),
                      lerpd(at_c(f, nx, ny, nz, i, j + 1, k), at_c(f, nx, ny, nz, i + 1, j + 1, k), cx), cy);
    double v2 = lerpd(lerpd(at_c(f, nx, ny, nz, i, j, k + 1), at_c(f, nx, ny, nz, i + 1, j, k + 1), cx),
                      lerpd(at_c(f, nx, ny, nz, i, j + 1, k + 1), at_c(f, nx, ny, nz, i + 1, j + 1, k + 1), cx),
                      cy);
    return (float)lerpd(v1, v2, cz);
}

// Cubic Lagrange weights on {-1,0,1,2}. Buffer3D<float>::cubic_interp takes the
// fraction as the buffer's element type, so the weights are evaluated in float
// and combined with the (double) values in double.
static inline double cubic_interp(double m1, double v0, double v1, double v2, double fd) {
    const float f = (float)fd;
    const float f2 = f * f, f3 = f2 * f;
    const float wm1 = -float(1. / 3) * f + float(1. / 2) * f2 - float(1. / 6) * f3;
    const float w0 = 1 - f2 + float(1. / 2) * (f3 - f);
    const float w1 = f + float(1. / 2) * (f2 - f3);
    const float w2 = float(1. / 6) * (f3 - f);
    return wm1 * m1 + w0 * v0 + w1 * v1 + w2 * v2;
}

float or_sample_cubic(const float* f, int nx, int ny, int nz, double h, double ox, double oy,
                      double oz, double wx, double wy, double wz) {
    double gx = wx / h + ox, gy = wy / h + oy, gz = wz / h + oz;
    int i = (int)std::floor(gx), j = (int)std::floor(gy), k = (int)std::floor(gz);
    double x = gx - (double)i, y = gy - (double)j, z = gz - (double)k;
    double fz[4];
    for (int dk = -1; dk <= 2; ++dk) {
        double fy[4];
        for (int dj = -1; dj <= 2; ++dj)
            fy[dj + 1] = cubic_interp(at_c(f, nx, ny, nz, i - 1, j + dj, k + dk), at_c(f, nx, ny, nz, i, j + dj, k + dk),
                                      at_c(f, nx, ny, nz, i + 1, j + dj, k + dk),
                                      at_c(f, nx, ny, nz, i + 2, j + dj, k + dk), x);
        fz[dk + 1] = cubic_interp(fy[0], fy[1], fy[2], fy[3], y);
    }
    return (float)cubic_interp(fz[0], fz[1], fz[2], fz[3], z);
}

float or_sample_cube_lerp(const float* f, int nx, int ny, int nz, double h, double ox, double oy,
                          double oz, double wx, double wy, double wz, float* corners) {
    double gx = wx / h + ox, gy = wy / h + oy, gz = wz / h + oz;
    int i = (int)std::floor(gx), j = (int)std::floor(gy), k = (int)std::floor(gz);
    for (int q = 0; q < 8; ++q)
        corners[q] = at_c(f, nx, ny, nz, i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1));
    return or_sample_linear(f, nx, ny, nz, h, ox, oy, oz, wx, wy, wz);
}

// ---------------------------------------------------------------------------
// Semi-Lagrangian advection, RK2 midpoint backtrace.
// mode 0: trilinear gather  -- SubDomainSolver::advect_field, S/sub_domain.cpp:191-235
// mode 1: clamped cubic     -- SmokeSolver3D::advect_field_cubic_clamp,
//                              S/Smoke_solver3D.cpp:657-707
// Velocity grids u (nx+1,ny,nz; offset .5,0,0), v, w as in S/Smoke_solver3D.h:299-307.
// field: fnx x fny x fnz with offsets (fox,foy,foz); output field_new same shape.
// ---------------------------------------------------------------------------
void or_advect_rk2(int nx, int ny, int nz, float h, float dt, int mode, const float* u,
                   const float* v, const float* w, const float* field, int fnx, int fny, int fnz,
                   double fox, double foy, double foz, float* out) {
    const double hd = h;
#pragma omp parallel for
    for (int k = 0; k < fnz; ++k)
        for (int j = 0; j < fny; ++j)
            for (int i = 0; i < fnx; ++i) {
                float wx = (float)(((double)(float)i - fox) * hd);
                float wy = (float)(((double)(float)j - foy) * hd);
                float wz = (float)(((double)(float)k - foz) * hd);
                float uu = or_sample_linear(u, nx + 1, ny, nz, hd, 0.5, 0, 0, wx, wy, wz);
                float vv = or_sample_linear(v, nx, ny + 1, nz, hd, 0, 0.5, 0, wx, wy, wz);
                float ww = or_sample_linear(w, nx, ny, nz + 1, hd, 0, 0, 0.5, wx, wy, wz);
                float px = (float)(wx - 0.5 * dt * uu), py = (float)(wy - 0.5 * dt * vv),
                      pz = (float)(wz - 0.5 * dt * ww);
                uu = or_sample_linear(u, nx + 1, ny, nz, hd, 0.5, 0, 0, px, py, pz);
                vv = or_sample_linear(v, nx, ny + 1, nz, hd, 0, 0.5, 0, px, py, pz);
                ww = or_sample_linear(w, nx, ny, nz + 1, hd, 0, 0, 0.5, px, py, pz);
                if (mode == 0) {
                    px = wx - dt * uu; py = wy - dt * vv; pz = wz - dt * ww;
                } else {
                    // advect_field_cubic_clamp: px = world_x - dt*u computed as float
                    px = wx - dt * uu; py = wy - dt * vv; pz = wz - dt * ww;
                }
                float val;
                if (mode == 0) {
                    val = or_sample_linear(field, fnx, fny, fnz, hd, fox, foy, foz, px, py, pz);
                } else {
                    float cn[8];
                    float lin_v = or_sample_cube_lerp(field, fnx, fny, fnz, hd, fox, foy, foz, px, py, pz, cn);
                    float cub = or_sample_cubic(field, fnx, fny, fnz, hd, fox, foy, foz, px, py, pz);
                    float mn = cn[0], mx = cn[0];
                    for (int q = 1; q < 8; ++q) { mn = std::min(mn, cn[q]); mx = std::max(mx, cn[q]); }
                    val = cub;
                    if (cub <= mn || cub >= mx) val = lin_v;
                }
                out[lin(i, j, k, fnx, fny)] = val;
            }
}

// ---------------------------------------------------------------------------
// Particle binning: cell = trunc((p - bmin + 0.5h)/h) per axis (float ops,
// truncating conversion, S/vec.h:37-39). clamp=0 drops out-of-range particles
// (coarse grid, S/Smoke_solver3D.cpp:1327-1346); clamp=1 clamps them into the
// grid (subdomain, S/sub_domain.cpp:1087-1107). Within a cell, particles keep
// their input order (append order of the reference's per-cell lists).
// Outputs: cell_start[ncell+1] (prefix offsets), order[...] particle ids; returns
// the number of binned particles. pos is xyz-interleaved float.
// ---------------------------------------------------------------------------
int or_hash(int np, const float* pos, float h, int nx, int ny, int nz, int clamp, float bx,
            float by, float bz, int* cell_key, int* cell_start, int* order) {
    const float half = 0.5f * h;  // Vec3f(0.5,0.5,0.5)*_hx
    const int ncell = nx * ny * nz;
    std::vector<int> cnt(ncell + 1, 0);
    for (int p = 0; p < np; ++p) {
        float x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
        if (clamp) { x = x - bx; y = y - by; z = z - bz; }
        x = x + half; y = y + half; z = z + half;
        int i = trunc_i32(x / h), j = trunc_i32(y / h), k = trunc_i32(z / h);
        int key = -1;
        if (clamp) {
            i = clampv(i, 0, nx - 1); j = clampv(j, 0, ny - 1); k = clampv(k, 0, nz - 1);
            key = k * nx * ny + j * nx + i;
        } else if (i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz) {
            key = k * nx * ny + j * nx + i;
        }
        cell_key[p] = key;
        if (key >= 0) cnt[key + 1]++;
    }
    for (int c = 0; c < ncell; ++c) cnt[c + 1] += cnt[c];
    std::memcpy(cell_start, cnt.data(), sizeof(int) * (ncell + 1));
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int p = 0; p < np; ++p)
        if (cell_key[p] >= 0) order[fill[cell_key[p]]++] = p;
    return cnt[ncell];
}

// Particle level set of the liquid (the oracle of vflip_liquid_phi): at every
// cell centre (i h, j h, k h), phi = min |centre - p| over the particles binned
// (as or_hash, clamp 0) in the 3x3x3 cells around it, minus radius; 3h - radius
// where no particle is that close. Distances in float, no contraction.
void or_liquid_phi(int np, const float* pos, float h, int nx, int ny, int nz, float radius, float* phi) {
    const int ncell = nx * ny * nz;
    for (int c = 0; c < ncell; ++c) phi[c] = 3.0f * h;
    const float half = 0.5f * h;
    for (int p = 0; p < np; ++p) {
        const float x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
        const int i = trunc_i32((x + half) / h), j = trunc_i32((y + half) / h), k = trunc_i32((z + half) / h);
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) continue;
        for (int ck = k - 1; ck <= k + 1; ++ck)
            for (int cj = j - 1; cj <= j + 1; ++cj)
                for (int ci = i - 1; ci <= i + 1; ++ci) {
                    if (ci < 0 || cj < 0 || ck < 0 || ci >= nx || cj >= ny || ck >= nz) continue;
                    const float dx = (float)ci * h - x, dy = (float)cj * h - y, dz = (float)ck * h - z;
                    const float d = std::sqrt(dx * dx + dy * dy + dz * dz);
                    float& f = phi[(ck * ny + cj) * nx + ci];
                    if (d < f) f = d;
                }
    }
    for (int c = 0; c < ncell; ++c) phi[c] = phi[c] - radius;
}

// ---------------------------------------------------------------------------
// P2G gather for one MAC component. Each face sample sums tent-weighted
// particle velocities over the (2l+1)^3 cells around its index and divides by
// (1e-5 + sum w). S/Smoke_solver3D.cpp:1451-1510 and S/Smoke_solver3D.h:211-229.
// Cells visited k-outer, i-inner; particles in binned order; float accumulation.
// field/coef: fnx x fny x fnz with offsets (fox,foy,foz) (u: nx+1,ny,nz,.5,0,0).
// sub=1 adds the subdomain's widened kernel near its boundary (l=4, support 4h
// within 4 cells of a face, S/sub_domain.h:89-103, S/sub_domain.cpp:1209-1265)
// and subtracts bmin from particle positions.
// ---------------------------------------------------------------------------
static inline float tent(float r) {
    float res = 0;
    if (r >= -1 && r < 0) res = 1 + r;
    if (r >= 0 && r < 1) res = 1 - r;
    return res;
}

void or_p2g(int np, const float* pos, const float* vel, int comp, float h, int nx, int ny, int nz,
            const int* cell_start, const int* order, int fnx, int fny, int fnz, double fox,
            double foy, double foz, int sub, float bx, float by, float bz, float* field,
            float* coef) {
    (void)np;
#pragma omp parallel for
    for (int k = 0; k < fnz; ++k)
        for (int j = 0; j < fny; ++j)
            for (int i = 0; i < fnx; ++i) {
                float gx = (float)(((double)(float)i - fox) * (double)h);
                float gy = (float)(((double)(float)j - foy) * (double)h);
                float gz = (float)(((double)(float)k - foz) * (double)h);
                float tw = 1e-5f, tv = 0;
                int l = 1;
                float hk = h;
                if (sub) {
                    int dist = std::min(std::min(std::min(i, nx - i), std::min(j, ny - j)), std::min(k, nz - k));
                    if (dist <= 4) { l = 4; hk = 4.0f * h; }
                }
                for (int dk = -l; dk <= l; ++dk)
                    for (int dj = -l; dj <= l; ++dj)
                        for (int di = -l; di <= l; ++di) {
                            int a = i + di, b = j + dj, c = k + dk;
                            if (a < 0 || a >= nx || b < 0 || b >= ny || c < 0 || c >= nz) continue;
                            int cell = c * nx * ny + b * nx + a;
                            for (int q = cell_start[cell]; q < cell_start[cell + 1]; ++q) {
                                int p = order[q];
                                float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
                                if (sub) { px = px - bx; py = py - by; pz = pz - bz; }
                                float wgt = tent((px - gx) / hk) * tent((py - gy) / hk) * tent((pz - gz) / hk);
                                tw += wgt;
                                tv += wgt * vel[3 * p + comp];
                            }
                        }
                size_t o = lin(i, j, k, fnx, fny);
                coef[o] = tw;
                field[o] = tv / tw;
            }
}

// ---------------------------------------------------------------------------
// G2P, pure FLIP: vel += trilinear(u_new - u_old) at the particle position.
// S/Smoke_solver3D.cpp:1298-1325 (getDelta) and 1398-1406 (particle_from_grid).
// du/dv/dw are the delta grids (MAC shapes). flip_ratio blends with PIC when < 1
// (PIC/FLIP mechanism of S/fluid_particle.h:297-336); flip_ratio=1 is the
// reference's pure FLIP path and ignores u/v/w.
// ---------------------------------------------------------------------------
void or_g2p(int np, const float* pos, float* vel, float h, int nx, int ny, int nz,
            const float* du, const float* dv, const float* dw, const float* u, const float* v,
            const float* w, float flip_ratio, float bx, float by, float bz) {
    const double hd = h;
#pragma omp parallel for
    for (int p = 0; p < np; ++p) {
        float x = pos[3 * p] - bx, y = pos[3 * p + 1] - by, z = pos[3 * p + 2] - bz;
        float a = or_sample_linear(du, nx + 1, ny, nz, hd, 0.5, 0, 0, x, y, z);
        float b = or_sample_linear(dv, nx, ny + 1, nz, hd, 0, 0.5, 0, x, y, z);
        float c = or_sample_linear(dw, nx, ny, nz + 1, hd, 0, 0, 0.5, x, y, z);
        if (flip_ratio >= 1.0f) {
            vel[3 * p] += a; vel[3 * p + 1] += b; vel[3 * p + 2] += c;
        } else {
            float pu = or_sample_linear(u, nx + 1, ny, nz, hd, 0.5, 0, 0, x, y, z);
            float pv = or_sample_linear(v, nx, ny + 1, nz, hd, 0, 0.5, 0, x, y, z);
            float pw = or_sample_linear(w, nx, ny, nz + 1, hd, 0, 0, 0.5, x, y, z);
            // pic_w*u_new + flip_w*(vel + du) with double weights, as the
            // 0.03/0.97 blend of S/fluid_particle.h:316-329
            const double fw = (double)flip_ratio, pw_ = 1.0 - fw;
            vel[3 * p] = (float)(pw_ * pu + fw * (vel[3 * p] + a));
            vel[3 * p + 1] = (float)(pw_ * pv + fw * (vel[3 * p + 1] + b));
            vel[3 * p + 2] = (float)(pw_ * pw + fw * (vel[3 * p + 2] + c));
        }
    }
}

// ---------------------------------------------------------------------------
// Coarse/fine operators. S/coarse_fine_solver.cpp:73-108 (fine->coarse box
// average with a face offset), 146-170 (coarse->fine trilinear), 182-230
// (flux-matching face correction). Ratio N, fine spacing hf.
// ---------------------------------------------------------------------------
void or_fine_to_coarse(const float* fine, int fnx, int fny, int fnz, float* coarse, int cnx,
                       int cny, int cnz, int N, int offx, int offy, int offz) {
#pragma omp parallel for
    for (int k = 0; k < cnz; ++k)
        for (int j = 0; j < cny; ++j)
            for (int i = 0; i < cnx; ++i) {
                int count = 0;
                float q = 0;
                for (int a = std::max(0, i * N - offx); a < std::min(fnx, i * N + N - offx); ++a)
                    for (int b = std::max(0, j * N - offy); b < std::min(fny, j * N + N - offy); ++b)
                        for (int c = std::max(0, k * N - offz); c < std::min(fnz, k * N + N - offz); ++c) {
                            ++count;
                            q += fine[lin(a, b, c, fnx, fny)];
                        }
                coarse[lin(i, j, k, cnx, cny)] = count > 0 ? q / (float)count : 0;
            }
}

void or_coarse_to_fine(float* fine, int fnx, int fny, int fnz, double fox, double foy, double foz,
                       float hf, const float* coarse, int cnx, int cny, int cnz, double cox,
                       double coy, double coz, double hc) {
#pragma omp parallel for
    for (int k = 0; k < fnz; ++k)
        for (int j = 0; j < fny; ++j)
            for (int i = 0; i < fnx; ++i) {
                float px = ((float)i - (float)fox) * hf, py = ((float)j - (float)foy) * hf,
                      pz = ((float)k - (float)foz) * hf;
                fine[lin(i, j, k, fnx, fny)] = or_sample_linear(coarse, cnx, cny, cnz, hc, cox, coy, coz, px, py, pz);
            }
}

void or_set_face_boundary(float* fine, int fnx, int fny, int fnz, const float* coarse, int cnx,
                          int cny, int cnz, int N, int offx, int offy, int offz) {
    // Each coarse cell owns a disjoint fine block, so the loop is write-disjoint.
#pragma omp parallel for
    for (int k = 0; k < cnz; ++k)
        for (int j = 0; j < cny; ++j)
            for (int i = 0; i < cnx; ++i) {
                float q = 0;
                for (int a = std::max(0, i * N); a < std::min(fnx, i * N + offx); ++a)
                    for (int b = std::max(0, j * N); b < std::min(fny, j * N + offy); ++b)
                        for (int c = std::max(0, k * N); c < std::min(fnz, k * N + offz); ++c) q += fine[lin(a, b, c, fnx, fny)];
                float diff = coarse[lin(i, j, k, cnx, cny)] - q / (float)(N * N);
                for (int a = std::max(0, i * N); a < std::min(fnx, i * N + offx); ++a)
                    for (int b = std::max(0, j * N); b < std::min(fny, j * N + offy); ++b)
                        for (int c = std::max(0, k * N); c < std::min(fnz, k * N + offz); ++c) fine[lin(a, b, c, fnx, fny)] += diff;
            }
}

// ---------------------------------------------------------------------------
// Storage helpers: the reference Buffer3D 8^3-brick layout.
// idx = ((K*bx*by + J*bx + I) << 9) + (kk << 6) + (jj << 3) + ii, S/fluid_buffer3D.h:186-202.
// ---------------------------------------------------------------------------
void or_to_blocked(const float* src, int nx, int ny, int nz, float* dst) {
    int bx = (nx + 7) / 8, by = (ny + 7) / 8, bz = (nz + 7) / 8;
    std::memset(dst, 0, sizeof(float) * (size_t)bx * by * bz * 512);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t idx = (((size_t)(k >> 3) * bx * by + (size_t)(j >> 3) * bx + (i >> 3)) << 9) +
                             ((k & 7) << 6) + ((j & 7) << 3) + (i & 7);
                dst[idx] = src[lin(i, j, k, nx, ny)];
            }
}

unsigned int or_randhash(unsigned int seed) {
    // S/util.h:238-246
    unsigned int i = (seed ^ 0xA3C59AC3u) * 2654435769u;
    i ^= (i >> 16);
    i *= 2654435769u;
    i ^= (i >> 16);
    i *= 2654435769u;
    return i;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"

extern "C" {
// ---------------------------------------------------------------------------
// Plume forcing of the synthetic smoke workload (not reference code: the
// workload re-specifies the reference's cylinder source on a 2x grid and its
// particle buoyancy, S/Smoke_solver3D.cpp:1624-1668, as a sphere source and
// face buoyancy). Source: rho = T = 1 + jitter*U where U = randhash(seed*0x9E3779B1
// + cell)/UINT_MAX inside |x - c| <= r (float ops); buoyancy on v faces 1..nj-1:
// v += dt*(-alpha*rho_f + beta*T_f) with face averages 0.5*(a + b).
// ---------------------------------------------------------------------------
void or_plume_force(int ni, int nj, int nk, float h, float dt, float cx, float cy, float cz, float r,
                    unsigned seed, float jitter, float alpha, float beta, float* rho, float* temp, float* v) {
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                float dx = i * h - cx, dy = j * h - cy, dz = k * h - cz;
                if (dx * dx + dy * dy + dz * dz > r * r) continue;
                size_t c = lin(i, j, k, ni, nj);
                float u = (float)or_randhash(seed * 0x9E3779B1u + (unsigned)c) * (1.0f / 4294967295.0f);
                rho[c] = 1.0f + jitter * u;
                temp[c] = 1.0f + jitter * u;
            }
    for (int k = 0; k < nk; ++k)
        for (int j = 1; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                size_t c = lin(i, j, k, ni, nj), cm = c - ni;
                float rf = 0.5f * (rho[c] + rho[cm]);
                float tf = 0.5f * (temp[c] + temp[cm]);
                v[lin(i, j, k, ni, nj + 1)] += dt * (-alpha * rf + beta * tf);
            }
}

// ---------------------------------------------------------------------------
// Buffer3D::sample_grad (S/fluid_buffer3D.h:249-258): central differences of
// sample_linear at +-h (the buffer's double spacing), each difference taken in
// float, divided by h in double, stored as float and scaled by 0.5f.
// ---------------------------------------------------------------------------
void or_sample_grad(const float* f, int nx, int ny, int nz, double h, double ox, double oy, double oz,
                    double wx, double wy, double wz, float* g) {
    const float right = or_sample_linear(f, nx, ny, nz, h, ox, oy, oz, wx + h, wy, wz);
    const float left = or_sample_linear(f, nx, ny, nz, h, ox, oy, oz, wx - h, wy, wz);
    const float up = or_sample_linear(f, nx, ny, nz, h, ox, oy, oz, wx, wy + h, wz);
    const float bottom = or_sample_linear(f, nx, ny, nz, h, ox, oy, oz, wx, wy - h, wz);
    const float front = or_sample_linear(f, nx, ny, nz, h, ox, oy, oz, wx, wy, wz - h);
    const float back = or_sample_linear(f, nx, ny, nz, h, ox, oy, oz, wx, wy, wz + h);
    g[0] = 0.5f * (float)((double)(right - left) / h);
    g[1] = 0.5f * (float)((double)(up - bottom) / h);
    g[2] = 0.5f * (float)((double)(back - front) / h);
}

// normalize(Vec3f): a /= mag(a), mag = sqrt(mag2) + 1e-6 (S/vec.h:194-203,
// 218-219). mag2 accumulates the squares in float; the reference's unqualified
// sqrt(float) binds to the double ::sqrt when its headers are compiled with
// g++/clang (checked against oracle/_ref), so the root is taken in double.
void or_normalize3(float* a) {
    float l = a[0] * a[0];
    l += a[1] * a[1];
    l += a[2] * a[2];
    const float m = (float)(std::sqrt((double)l) + 1e-6);
    a[0] /= m;
    a[1] /= m;
    a[2] /= m;
}

// ---------------------------------------------------------------------------
// FLIP particle advection: traceRK3 then the solid push-out of advect_FLIP.
// Coarse grid (sub=0): S/Smoke_solver3D.cpp:141-154 (traceRK3, where
// c3 = (1 - c1 - c2)*dt with c1, c2 already scaled by dt) and 1408-1430
// (advect_FLIP: if phi(pos) < 0, pos -= phi*normalize(grad phi)).
// Subdomain (sub=1): S/sub_domain.cpp:116-129 (c3 = 4/9 dt; velocity sampled at
// pos - bmin) and 1180-1202 (phi and its gradient sampled at the unshifted pos).
// Velocity: trilinear u (nx+1,ny,nz; .5,0,0), v, w (S/Smoke_solver3D.cpp:71-88,
// without the subdomain blend). phi: nodal (nx+1,ny+1,nz+1; .5,.5,.5).
// pos is xyz-interleaved float, updated in place.
// ---------------------------------------------------------------------------
void or_advect_particles(int np, float* pos, const float* u, const float* v, const float* w,
                         const float* phi, int nx, int ny, int nz, float h, float dt, int sub,
                         float bx, float by, float bz) {
    const double hd = h;
    const float c1 = (float)(2.0 / 9.0 * dt), c2 = (float)(3.0 / 9.0 * dt);
    const float c3 = sub ? (float)(4.0 / 9.0 * dt) : (float)((1.0 - c1 - c2) * dt);
    const float t1 = (float)(0.5 * dt), t2 = (float)(0.75 * dt);
    auto vel = [&](const float* x, float* o) {
        float px = x[0], py = x[1], pz = x[2];
        if (sub) { px = px - bx; py = py - by; pz = pz - bz; }
        o[0] = or_sample_linear(u, nx + 1, ny, nz, hd, 0.5, 0, 0, px, py, pz);
        o[1] = or_sample_linear(v, nx, ny + 1, nz, hd, 0, 0.5, 0, px, py, pz);
        o[2] = or_sample_linear(w, nx, ny, nz + 1, hd, 0, 0, 0.5, px, py, pz);
    };
#pragma omp parallel for
    for (int p = 0; p < np; ++p) {
        float x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        float v1[3], v2[3], v3[3], m1[3], m2[3];
        vel(x, v1);
        for (int a = 0; a < 3; ++a) m1[a] = x[a] + t1 * v1[a];
        vel(m1, v2);
        for (int a = 0; a < 3; ++a) m2[a] = x[a] + t2 * v2[a];
        vel(m2, v3);
        for (int a = 0; a < 3; ++a) x[a] = x[a] + c1 * v1[a] + c2 * v2[a] + c3 * v3[a];
        const float ph = or_sample_linear(phi, nx + 1, ny + 1, nz + 1, hd, 0.5, 0.5, 0.5, x[0], x[1], x[2]);
        if (ph < 0) {
            float g[3];
            or_sample_grad(phi, nx + 1, ny + 1, nz + 1, hd, 0.5, 0.5, 0.5, x[0], x[1], x[2], g);
            or_normalize3(g);
            for (int a = 0; a < 3; ++a) x[a] -= ph * g[a];
        }
        pos[3 * p] = x[0];
        pos[3 * p + 1] = x[1];
        pos[3 * p + 2] = x[2];
    }
}

// getDelta (S/Smoke_solver3D.cpp:1297-1318): out = new - old, elementwise float.
void or_get_delta(int n, const float* fnew, const float* fold, float* out) {
    for (int c = 0; c < n; ++c) out[c] = fnew[c] - fold[c];
}

// frand(-1, 1) of the reseeding (S/Smoke_solver3D.h:36-39: a + (b - a) U) with
// U drawn from the counter hash instead of rand(): U = (randhash(seed *
// 0x9E3779B1 + key) >> 8) / 2^24, key = 3 (8 cell + slot) + axis.
float or_reseed_jitter(unsigned seed, unsigned key) {
    const unsigned r = or_randhash(seed * 0x9E3779B1u + key);
    const float u = (float)(r >> 8) * (1.0f / 16777216.0f);
    return -1.0f + 2.0f * u;
}

// reorder_FLIP. sub = 0: SmokeSolver3D, S/Smoke_solver3D.cpp:1347-1396 (every
// cell; new particle at (i,j,k) h + 0.49 h r; velocity get_velocity_self and
// solid phi at that position). sub = 1: SubDomainSolver,
// S/sub_domain.cpp:1110-1159 (interior cells 1..n-2 only, the boundary ring's
// particles are dropped; new particle at bmin + (i,j,k) h + 0.49 h r; velocity
// get_velocity and solid phi at pos - bmin, S/sub_domain.cpp:70-90).
// Cell-major over the hash lists (cell_start/order from or_hash): a cell with c
// particles keeps them all when c < 10, else its first 8; when c < 8 it also
// gets 8 - c new ones. Then the particles whose trilinear nodal solid phi is
// < 0 are dropped, keeping the order. Writes min(count, cap) particles and
// returns the count.
int or_reorder(int nx, int ny, int nz, float h, int sub, float bx, float by, float bz, unsigned seed,
               const float* pos, const float* vel, const int* cell_start, const int* order, const float* u,
               const float* v, const float* w, const float* phi, int cap, float* out_pos, float* out_vel) {
    const double hd = h;
    const float b[3] = {bx, by, bz};
    auto solid_ok = [&](const float* x) {
        float q[3] = {x[0], x[1], x[2]};
        if (sub) for (int a = 0; a < 3; ++a) q[a] = x[a] - b[a];
        return or_sample_linear(phi, nx + 1, ny + 1, nz + 1, hd, 0.5, 0.5, 0.5, q[0], q[1], q[2]) >= 0.0f;
    };
    int out = 0;
    auto push = [&](const float* x, const float* vv) {
        if (!solid_ok(x)) return;
        if (out < cap)
            for (int a = 0; a < 3; ++a) {
                out_pos[3 * (size_t)out + a] = x[a];
                out_vel[3 * (size_t)out + a] = vv[a];
            }
        ++out;
    };
    const int lo = sub ? 1 : 0;
    for (int k = lo; k < nz - lo; ++k)
        for (int j = lo; j < ny - lo; ++j)
            for (int i = lo; i < nx - lo; ++i) {
                const size_t c = lin(i, j, k, nx, ny);
                const int s = cell_start[c], n = cell_start[c + 1] - s;
                const int keep = n < 10 ? n : 8;
                for (int q = 0; q < keep; ++q) {
                    const int p = order[s + q];
                    push(pos + 3 * (size_t)p, vel + 3 * (size_t)p);
                }
                const int ijk[3] = {i, j, k};
                for (int t = 0; t < 8 - n; ++t) {
                    float x[3], vv[3];
                    for (int a = 0; a < 3; ++a) {
                        const float r = or_reseed_jitter(seed, (unsigned)(3 * (8 * c + (size_t)t) + a));
                        const float cell = (float)ijk[a] * h;
                        const float base = sub ? b[a] + cell : cell;
                        x[a] = base + (r * 0.49f) * h;
                    }
                    float q[3] = {x[0], x[1], x[2]};
                    if (sub) for (int a = 0; a < 3; ++a) q[a] = x[a] - b[a];
                    vv[0] = or_sample_linear(u, nx + 1, ny, nz, hd, 0.5, 0, 0, q[0], q[1], q[2]);
                    vv[1] = or_sample_linear(v, nx, ny + 1, nz, hd, 0, 0.5, 0, q[0], q[1], q[2]);
                    vv[2] = or_sample_linear(w, nx, ny, nz + 1, hd, 0, 0, 0.5, q[0], q[1], q[2]);
                    push(x, vv);
                }
            }
    return out;
}

// ---------------------------------------------------------------------------
// Nested FLIP coupling: the coarse solver's view of one subdomain.
// OrCoarse = SmokeSolver3D's grids (nx,ny,nz cells of spacing h; MAC u, v, w;
// nodal solid phi; MAC solid velocity us, vs, ws, NULL = zero). OrSub =
// SubDomainSolver's members (cells, _hx, bmin, bmax, the same fields); its
// samplers take pos - bmin (S/sub_domain.cpp:60-95).
// ---------------------------------------------------------------------------
struct OrCoarse {
    int nx, ny, nz;
    float h;
    const float *u, *v, *w, *phi, *us, *vs, *ws;
};
struct OrSub {
    int nx, ny, nz;
    float h;
    float bmin[3], bmax[3];
    const float *u, *v, *w, *phi, *us, *vs, *ws;
};

static void or_mac_sample(const float* u, const float* v, const float* w, int nx, int ny, int nz, float h,
                          const float* x, float* o) {
    const double hd = h;
    o[0] = u ? or_sample_linear(u, nx + 1, ny, nz, hd, 0.5, 0, 0, x[0], x[1], x[2]) : 0.f;
    o[1] = v ? or_sample_linear(v, nx, ny + 1, nz, hd, 0, 0.5, 0, x[0], x[1], x[2]) : 0.f;
    o[2] = w ? or_sample_linear(w, nx, ny, nz + 1, hd, 0, 0, 0.5, x[0], x[1], x[2]) : 0.f;
}
static void or_sub_local(const OrSub& S, const float* x, float* q) {
    for (int a = 0; a < 3; ++a) q[a] = x[a] - S.bmin[a];
}
// is_in (margin 2) / is_in2 (margin 5), S/sub_domain.h:177-201: bmin + m*_hx in double.
static bool or_sub_inside(const OrSub& S, const float* x, double m) {
    for (int a = 0; a < 3; ++a)
        if (!(x[a] > S.bmin[a] + m * S.h && x[a] < S.bmax[a] - m * S.h)) return false;
    return true;
}
// get_boundary_coef, S/sub_domain.cpp:96-105.
float or_boundary_coef(const OrSub& S, const float* x) {
    const float e = 5.0f * S.h;
    float d = 0;
    for (int side = 0; side < 2; ++side) {
        float m = 0;
        for (int a = 0; a < 3; ++a) {
            const float b = side == 0 ? S.bmin[a] + e : S.bmax[a] - e;
            const float t = std::fabs(x[a] - b);
            if (a == 0 || t < m) m = t;
        }
        d = side == 0 ? m : (m < d ? m : d);
    }
    const float c = d / (4.0f * S.h);
    return c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
}
// SmokeSolver3D::get_velocity, S/Smoke_solver3D.cpp:71-88.
static void or_nested_velocity(const OrCoarse& C, const OrSub& S, const float* x, float* o) {
    or_mac_sample(C.u, C.v, C.w, C.nx, C.ny, C.nz, C.h, x, o);
    if (or_sub_inside(S, x, 5.0)) {
        float q[3], o2[3];
        or_sub_local(S, x, q);
        or_mac_sample(S.u, S.v, S.w, S.nx, S.ny, S.nz, S.h, q, o2);
        const float c = or_boundary_coef(S, x);
        for (int a = 0; a < 3; ++a) o[a] = c * o2[a] + (1.0f - c) * o[a];
    }
}
// get_solid_phi / get_solid_normal / get_solid_vel, S/Smoke_solver3D.cpp:60-124.
static float or_nested_phi(const OrCoarse& C, const OrSub& S, const float* x) {
    if (or_sub_inside(S, x, 2.0)) {
        float q[3];
        or_sub_local(S, x, q);
        return or_sample_linear(S.phi, S.nx + 1, S.ny + 1, S.nz + 1, S.h, 0.5, 0.5, 0.5, q[0], q[1], q[2]);
    }
    return or_sample_linear(C.phi, C.nx + 1, C.ny + 1, C.nz + 1, C.h, 0.5, 0.5, 0.5, x[0], x[1], x[2]);
}
static void or_nested_normal(const OrCoarse& C, const OrSub& S, const float* x, float* g) {
    if (or_sub_inside(S, x, 2.0)) {
        float q[3];
        or_sub_local(S, x, q);
        or_sample_grad(S.phi, S.nx + 1, S.ny + 1, S.nz + 1, S.h, 0.5, 0.5, 0.5, q[0], q[1], q[2], g);
    } else {
        or_sample_grad(C.phi, C.nx + 1, C.ny + 1, C.nz + 1, C.h, 0.5, 0.5, 0.5, x[0], x[1], x[2], g);
    }
    or_normalize3(g);
}
static void or_nested_solid_vel(const OrCoarse& C, const OrSub& S, const float* x, float* o) {
    if (or_sub_inside(S, x, 2.0)) {
        float q[3];
        or_sub_local(S, x, q);
        or_mac_sample(S.us, S.vs, S.ws, S.nx, S.ny, S.nz, S.h, q, o);
    } else {
        or_mac_sample(C.us, C.vs, C.ws, C.nx, C.ny, C.nz, C.h, x, o);
    }
}

void or_nested_velocity_at(const OrCoarse* C, const OrSub* S, int n, const float* pos, float* out) {
    for (int p = 0; p < n; ++p) or_nested_velocity(*C, *S, pos + 3 * (size_t)p, out + 3 * (size_t)p);
}

// advect_FLIP of the coarse particles with a subdomain: traceRK3 through the
// blended get_velocity, push-out with the subdomain's phi and normal inside
// is_in (S/Smoke_solver3D.cpp:141-154, 1408-1430).
void or_advect_particles_nested(const OrCoarse* Cp, const OrSub* Sp, float dt, int np, float* pos) {
    const OrCoarse& C = *Cp;
    const OrSub& S = *Sp;
    const float c1 = (float)(2.0 / 9.0 * dt), c2 = (float)(3.0 / 9.0 * dt);
    const float c3 = (float)((1.0 - c1 - c2) * dt);
    const float t1 = (float)(0.5 * dt), t2 = (float)(0.75 * dt);
#pragma omp parallel for
    for (int p = 0; p < np; ++p) {
        float x[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        float v1[3], v2[3], v3[3], m1[3], m2[3];
        or_nested_velocity(C, S, x, v1);
        for (int a = 0; a < 3; ++a) m1[a] = x[a] + t1 * v1[a];
        or_nested_velocity(C, S, m1, v2);
        for (int a = 0; a < 3; ++a) m2[a] = x[a] + t2 * v2[a];
        or_nested_velocity(C, S, m2, v3);
        for (int a = 0; a < 3; ++a) x[a] = x[a] + c1 * v1[a] + c2 * v2[a] + c3 * v3[a];
        const float ph = or_nested_phi(C, S, x);
        if (ph < 0) {
            float g[3];
            or_nested_normal(C, S, x, g);
            for (int a = 0; a < 3; ++a) x[a] -= ph * g[a];
        }
        for (int a = 0; a < 3; ++a) pos[3 * p + a] = x[a];
    }
}

// The halo copy of S/Smoke_solver3D.cpp:393-401: particles with is_halo(pos, dx)
// (S/sub_domain.h:203-214: bmin - dx < pos < bmax + dx in float), in order.
int or_halo_select(const OrSub* Sp, float dx, int np, const float* pos, const float* vel, int cap, float* out_pos,
                   float* out_vel) {
    const OrSub& S = *Sp;
    int n = 0;
    for (int p = 0; p < np; ++p) {
        bool in = true;
        for (int a = 0; a < 3; ++a) {
            const float x = pos[3 * (size_t)p + a];
            in = in && x > S.bmin[a] - dx && x < S.bmax[a] + dx;
        }
        if (!in) continue;
        if (n < cap)
            for (int a = 0; a < 3; ++a) {
                out_pos[3 * (size_t)n + a] = pos[3 * (size_t)p + a];
                out_vel[3 * (size_t)n + a] = vel[3 * (size_t)p + a];
            }
        ++n;
    }
    return n;
}

// seed_ghost, S/Smoke_solver3D.cpp:1529-1560: for every subdomain cell whose
// (coarse) get_solid_phi at bmin + (i,j,k) h is < 3h, 8 candidates at
// (i,j,k) h + 0.49 h r + bmin (r = frand(-1,1) per axis, here the counter hash
// of or_reseed_jitter with key 3 (8 cell + slot) + axis); a candidate whose
// subdomain phi lies in (-2h, 2h] is kept with velocity get_velocity (blended),
// or get_solid_vel where that phi is <= 0. Cell-major order. Writes min(n, cap)
// and returns n.
int or_seed_ghost(const OrCoarse* Cp, const OrSub* Sp, unsigned seed, int cap, float* out_pos, float* out_vel) {
    const OrCoarse& C = *Cp;
    const OrSub& S = *Sp;
    const float h = S.h;
    int n = 0;
    for (int k = 0; k < S.nz; ++k)
        for (int j = 0; j < S.ny; ++j)
            for (int i = 0; i < S.nx; ++i) {
                const int ijk[3] = {i, j, k};
                float cpos[3];
                for (int a = 0; a < 3; ++a) cpos[a] = (float)ijk[a] * h + S.bmin[a];
                if (!(or_nested_phi(C, S, cpos) < 3.0f * h)) continue;
                const size_t c = lin(i, j, k, S.nx, S.ny);
                for (int r = 0; r < 8; ++r) {
                    float x[3], q[3], vv[3];
                    for (int a = 0; a < 3; ++a) {
                        const float jit = or_reseed_jitter(seed, (unsigned)(3 * (8 * c + (size_t)r) + a));
                        x[a] = ((float)ijk[a] * h + (jit * 0.49f) * h) + S.bmin[a];
                    }
                    or_sub_local(S, x, q);
                    const float phi =
                        or_sample_linear(S.phi, S.nx + 1, S.ny + 1, S.nz + 1, h, 0.5, 0.5, 0.5, q[0], q[1], q[2]);
                    if (!(phi <= 2.0 * h && phi > -2.0 * h)) continue;
                    or_nested_velocity(C, S, x, vv);
                    if (phi <= 0) or_nested_solid_vel(C, S, x, vv);
                    if (n < cap)
                        for (int a = 0; a < 3; ++a) {
                            out_pos[3 * (size_t)n + a] = x[a];
                            out_vel[3 * (size_t)n + a] = vv[a];
                        }
                    ++n;
                }
            }
    return n;
}

// PSE_kernel, S/Smoke_solver3D.h:204-209 (as written: 1/frac*frac*frac).
float or_pse_kernel(const float* xi, const float* xj, float tau) {
    const float frac = (float)std::sqrt(4.0 * 3.1415926 * tau);
    float d = (xi[0] - xj[0]) * (xi[0] - xj[0]);
    d += (xi[1] - xj[1]) * (xi[1] - xj[1]);
    d += (xi[2] - xj[2]) * (xi[2] - xj[2]);
    return (float)((double)(1.0f / frac * frac * frac) * std::exp(-d * 0.25 / tau));
}

// diffuse, S/Smoke_solver3D.cpp:1562-1621: for each of the first np_real
// subdomain particles (the ghosts follow them in pos/vel), when the coarse
// get_solid_phi at it is <= 2 h_coarse and its cell (trunc((x - bmin)/h + 0.5))
// lies in the subdomain: du = sum_j (v_j - v_i) K_ij / (1e-9 + sum_j K_ij) over
// the particles binned (cell_start/order: or_hash clamp 1 over all particles)
// in the 3x3x3 cells around it, z, y, x ascending, tau = dt nu + 1e-10.
// out_vel = vel + du for the first np_real particles.
void or_pse_diffuse(const OrCoarse* Cp, const OrSub* Sp, float dt, float nu, int np_real, const float* pos,
                    const float* vel, const int* cell_start, const int* order, float* out_vel) {
    const OrCoarse& C = *Cp;
    const OrSub& S = *Sp;
    const float tau = (float)(dt * nu + 1e-10);
#pragma omp parallel for
    for (int i = 0; i < np_real; ++i) {
        const float* xi = pos + 3 * (size_t)i;
        const float* ui = vel + 3 * (size_t)i;
        int ijk[3];
        for (int a = 0; a < 3; ++a) ijk[a] = trunc_i32((xi[a] - S.bmin[a]) / S.h + 0.5f);
        float du[3] = {0.f, 0.f, 0.f};
        if (or_nested_phi(C, S, xi) <= 2.0 * C.h && ijk[0] >= 0 && ijk[0] < S.nx && ijk[1] >= 0 && ijk[1] < S.ny &&
            ijk[2] >= 0 && ijk[2] < S.nz) {
            float coef = 1e-9f;
            float ue[3] = {0.f, 0.f, 0.f};
            for (int kk = std::max(ijk[2] - 1, 0); kk <= std::min(ijk[2] + 1, S.nz - 1); ++kk)
                for (int jj = std::max(ijk[1] - 1, 0); jj <= std::min(ijk[1] + 1, S.ny - 1); ++jj)
                    for (int ii = std::max(ijk[0] - 1, 0); ii <= std::min(ijk[0] + 1, S.nx - 1); ++ii) {
                        const size_t cidx = lin(ii, jj, kk, S.nx, S.ny);
                        for (int q = cell_start[cidx]; q < cell_start[cidx + 1]; ++q) {
                            const size_t pj = (size_t)order[q];
                            const float K = or_pse_kernel(xi, pos + 3 * pj, tau);
                            for (int a = 0; a < 3; ++a) ue[a] += (vel[3 * pj + a] - ui[a]) * K;
                            coef += K;
                        }
                    }
            for (int a = 0; a < 3; ++a) du[a] = ue[a] / coef;
        }
        for (int a = 0; a < 3; ++a) out_vel[3 * (size_t)i + a] = ui[a] + du[a];
    }
}

// ---------------------------------------------------------------------------
// Tracers: Vec4f (x, y, z, density) per tracer, S/Smoke_solver3D.h:45-161.
// With Sp == NULL the coarse grids alone are sampled (no subdomain).
// ---------------------------------------------------------------------------
static OrSub or_no_sub() {
    OrSub S{};
    for (int a = 0; a < 3; ++a) {
        S.bmin[a] = INFINITY;
        S.bmax[a] = -INFINITY;
    }
    S.h = 1.0f;
    return S;
}

// advect_tracers, S/Smoke_solver3D.h:101-125: traceRK3 (coarse c3), push-out
// where get_solid_phi <= 0, density d / (1 + decay dt).
void or_advect_tracers(const OrCoarse* Cp, const OrSub* Sp, float dt, float decay, int n, float* tr) {
    const OrCoarse& C = *Cp;
    const OrSub S = Sp ? *Sp : or_no_sub();
    const float c1 = (float)(2.0 / 9.0 * dt), c2 = (float)(3.0 / 9.0 * dt);
    const float c3 = (float)((1.0 - c1 - c2) * dt);
    const float t1 = (float)(0.5 * dt), t2 = (float)(0.75 * dt);
#pragma omp parallel for
    for (int p = 0; p < n; ++p) {
        float x[3] = {tr[4 * p], tr[4 * p + 1], tr[4 * p + 2]};
        float v1[3], v2[3], v3[3], m1[3], m2[3];
        or_nested_velocity(C, S, x, v1);
        for (int a = 0; a < 3; ++a) m1[a] = x[a] + t1 * v1[a];
        or_nested_velocity(C, S, m1, v2);
        for (int a = 0; a < 3; ++a) m2[a] = x[a] + t2 * v2[a];
        or_nested_velocity(C, S, m2, v3);
        for (int a = 0; a < 3; ++a) x[a] = x[a] + c1 * v1[a] + c2 * v2[a] + c3 * v3[a];
        const float ph = or_nested_phi(C, S, x);
        if (ph <= 0) {
            float g[3];
            or_nested_normal(C, S, x, g);
            for (int a = 0; a < 3; ++a) x[a] -= ph * g[a];
        }
        const float d = (float)(tr[4 * p + 3] / (1.0 + decay * dt));
        tr[4 * p] = x[0];
        tr[4 * p + 1] = x[1];
        tr[4 * p + 2] = x[2];
        tr[4 * p + 3] = d;
    }
}

// frand(a, b) = a + (b - a) U with U from the counter hash (as or_reseed_jitter).
static float or_frand(unsigned seed, unsigned key, float a, float b) {
    const unsigned r = or_randhash(seed * 0x9E3779B1u + key);
    const float u = (float)(r >> 8) * (1.0f / 16777216.0f);
    return a + (b - a) * u;
}

// The sphere emitter of emit_tracers, S/Smoke_solver3D.h:67-82: candidates
// (x, y, z) = frand(-r - hx, r + hx) per axis (key 3 t + axis for candidate t)
// accepted when x^2 + z^2 + y^2 <= r^2, the first n_emit accepted in candidate
// order, at centre + (x, y, z) with density 1. out[4 * n_emit].
void or_emit_sphere(unsigned seed, int n_emit, float cx, float cy, float cz, float r, float hx, float* out) {
    int got = 0;
    for (unsigned t = 0; got < n_emit; ++t) {
        const float x = or_frand(seed, 3 * t, -r - hx, r + hx);
        const float y = or_frand(seed, 3 * t + 1, -r - hx, r + hx);
        const float z = or_frand(seed, 3 * t + 2, -r - hx, r + hx);
        if (x * x + z * z + y * y <= r * r) {
            out[4 * (size_t)got] = cx + x;
            out[4 * (size_t)got + 1] = cy + y;
            out[4 * (size_t)got + 2] = cz + z;
            out[4 * (size_t)got + 3] = 1.0f;
            ++got;
        }
    }
}

// The subdomain part of emit_tracers, S/Smoke_solver3D.h:84-96: per subdomain
// cell (cell-major) 4 candidates bmin + (i,j,k) h + U h (U = frand(0, 1) per
// axis, key 3 (4 cell + slot) + axis), kept where 0 < get_solid_phi < h.
// Writes min(n, cap) float4 tracers and returns n.
int or_emit_sub_tracers(const OrCoarse* Cp, const OrSub* Sp, unsigned seed, int cap, float* out) {
    const OrCoarse& C = *Cp;
    const OrSub& S = *Sp;
    int n = 0;
    for (int k = 0; k < S.nz; ++k)
        for (int j = 0; j < S.ny; ++j)
            for (int i = 0; i < S.nx; ++i) {
                const int ijk[3] = {i, j, k};
                const size_t c = lin(i, j, k, S.nx, S.ny);
                for (int q = 0; q < 4; ++q) {
                    float x[3];
                    for (int a = 0; a < 3; ++a) {
                        const float u = or_frand(seed, (unsigned)(3 * (4 * c + (size_t)q) + a), 0.0f, 1.0f);
                        x[a] = (S.bmin[a] + (float)ijk[a] * S.h) + u * S.h;
                    }
                    const float ph = or_nested_phi(C, S, x);
                    if (!(ph > 0 && ph < S.h)) continue;
                    if (n < cap) {
                        for (int a = 0; a < 3; ++a) out[4 * (size_t)n + a] = x[a];
                        out[4 * (size_t)n + 3] = 1.0f;
                    }
                    ++n;
                }
            }
    return n;
}

// write_tracers' record, S/Smoke_solver3D.h:126-148: (x/lx - 0.5, y/lx, z/lx - 0.5, d).
void or_tracer_frame(int n, const float* tr, float lx, float* out) {
    for (int p = 0; p < n; ++p) {
        out[4 * (size_t)p] = (float)(tr[4 * (size_t)p] / lx - 0.5);
        out[4 * (size_t)p + 1] = tr[4 * (size_t)p + 1] / lx;
        out[4 * (size_t)p + 2] = (float)(tr[4 * (size_t)p + 2] / lx - 0.5);
        out[4 * (size_t)p + 3] = tr[4 * (size_t)p + 3];
    }
}
}  // extern "C"

// ---------------------------------------------------------------------------
// Geometric multigrid, MultiGridSolver3D<T> (S/Multigrid3D.h), T = float or
// double, on linear x-fastest arrays. A read past a grid that the reference
// makes into its Buffer3D brick padding (zeroed, never written) reads 0 here.
namespace mg3d {

struct Dims {
    int x, y, z;
    size_t n() const { return (size_t)x * y * z; }
};

// MultiGrid::m_ComputeLevels + m_AssignMemoryEachLevel (S/Multigrid3D.h:100-137)
std::vector<Dims> levels(int gx, int gy, int gz) {
    int level = 0, x = gx, y = gy, z = gz;
    while ((size_t)x * y * z > 512) {
        level = level + 1;
        x = x / 2; if (x == 0) x = 1;
        y = y / 2; if (y == 0) y = 1;
        z = z / 2; if (z == 0) z = 1;
    }
    std::vector<Dims> s(level + 1);
    s[0] = {gx, gy, gz};
    for (int i = 1; i <= level; ++i)
        s[i] = {(s[i - 1].x / 2) == 0 ? 1 : s[i - 1].x / 2, (s[i - 1].y / 2) == 0 ? 1 : s[i - 1].y / 2,
                (s[i - 1].z / 2) == 0 ? 1 : s[i - 1].z / 2};
    return s;
}

// The six neighbour values of (i,j,k): indexes clamped into the grid, zeroed at
// the walls with Dirichlet (S/Multigrid3D.h:1040-1062).
template <class T>
inline void neighbours(const Dims& g, bool dir, const T* x, unsigned i, unsigned j, unsigned k, T& lv, T& rv, T& tv,
                       T& dv, T& fv, T& bv) {
    unsigned li = (i == 0) ? i : i - 1;
    unsigned ri = (i == (unsigned)g.x - 1) ? i : i + 1;
    unsigned ti = (j == (unsigned)g.y - 1) ? j : j + 1;
    unsigned di = (j == 0) ? j : j - 1;
    unsigned fi = (k == 0) ? k : k - 1;
    unsigned bi = (k == (unsigned)g.z - 1) ? k : k + 1;
    lv = x[lin(li, j, k, g.x, g.y)];
    rv = x[lin(ri, j, k, g.x, g.y)];
    tv = x[lin(i, ti, k, g.x, g.y)];
    dv = x[lin(i, di, k, g.x, g.y)];
    fv = x[lin(i, j, fi, g.x, g.y)];
    bv = x[lin(i, j, bi, g.x, g.y)];
    if (dir) {
        if (i == 0) lv = 0;
        if (i == (unsigned)g.x - 1) rv = 0;
        if (j == (unsigned)g.y - 1) tv = 0;
        if (j == 0) dv = 0;
        if (k == 0) fv = 0;
        if (k == (unsigned)g.z - 1) bv = 0;
    }
}

// m_Red_Black_Gauss (S/Multigrid3D.h:840-1129): the parity-1 pass, then the
// parity-0 pass, each in place.
template <class T>
void rbgs(const Dims& g, bool dir, const T* b, T* x, int iter_time, double omega) {
    for (int it = 0; it < iter_time; ++it)
        for (int parity = 1; parity >= 0; --parity)
            for (unsigned k = 0; k < (unsigned)g.z; ++k)
                for (unsigned j = 0; j < (unsigned)g.y; ++j)
                    for (unsigned i = 0; i < (unsigned)g.x; ++i) {
                        if ((int)((j + k + i) % 2) != parity) continue;
                        const size_t c = lin(i, j, k, g.x, g.y);
                        T rhs = b[c];
                        T lv, rv, tv, dv, fv, bv;
                        neighbours(g, dir, x, i, j, k, lv, rv, tv, dv, fv, bv);
                        T rr = lv + rv;
                        rr += tv + dv;
                        rr += fv + bv;
                        rr -= rhs;
                        x[c] = (1 - omega) * x[c] + omega * rr / 6.0;
                    }
}

// m_ComputeResidual (S/Multigrid3D.h:1130-1186)
template <class T> void residual(const Dims& g, bool dir, T* res, const T* b, const T* x) {
    for (unsigned k = 0; k < (unsigned)g.z; ++k)
        for (unsigned j = 0; j < (unsigned)g.y; ++j)
            for (unsigned i = 0; i < (unsigned)g.x; ++i) {
                const size_t c = lin(i, j, k, g.x, g.y);
                T rhs = b[c];
                T center = x[c];
                T lv, rv, tv, dv, fv, bv;
                neighbours(g, dir, x, i, j, k, lv, rv, tv, dv, fv, bv);
                T rr = lv + rv;
                rr += tv + dv;
                rr += fv + bv;
                rr -= 6 * center;
                res[c] = rhs - rr;
            }
}

template <class T> inline T read0(const Dims& g, const T* a, unsigned i, unsigned j, unsigned k) {
    return (i < (unsigned)g.x && j < (unsigned)g.y && k < (unsigned)g.z) ? a[lin(i, j, k, g.x, g.y)] : T(0);
}

// m_Restrict (S/Multigrid3D.h:1193-1245)
template <class T> void restrict_(const Dims& cg, const Dims& fg, const T* src, T* dst) {
    for (unsigned k = 0; k < (unsigned)cg.z; ++k)
        for (unsigned j = 0; j < (unsigned)cg.y; ++j)
            for (unsigned i = 0; i < (unsigned)cg.x; ++i) {
                unsigned i_fine = i * 2, j_fine = j * 2, k_fine = k * 2;
                T r = 0;
                r += read0(fg, src, i_fine, j_fine, k_fine);
                r += read0(fg, src, i_fine + 1, j_fine, k_fine);
                r += read0(fg, src, i_fine, j_fine + 1, k_fine);
                r += read0(fg, src, i_fine + 1, j_fine + 1, k_fine);
                r += read0(fg, src, i_fine, j_fine, k_fine + 1);
                r += read0(fg, src, i_fine + 1, j_fine, k_fine + 1);
                r += read0(fg, src, i_fine, j_fine + 1, k_fine + 1);
                r += read0(fg, src, i_fine + 1, j_fine + 1, k_fine + 1);
                r *= 0.5;
                dst[lin(i, j, k, cg.x, cg.y)] = r;
            }
}

// m_Prolongate (S/Multigrid3D.h:1298-1326)
template <class T> void prolongate(const Dims& fg, const Dims& cg, const T* src, T* dst) {
    for (unsigned k = 0; k < (unsigned)fg.z; ++k)
        for (unsigned j = 0; j < (unsigned)fg.y; ++j)
            for (unsigned i = 0; i < (unsigned)fg.x; ++i) {
                T r = read0(cg, src, i / 2, j / 2, k / 2);
                dst[lin(i, j, k, fg.x, fg.y)] += r;
            }
}

template <class T> struct Solver {
    std::vector<Dims> sys;
    bool dir;
    std::vector<std::vector<T>> xs, bs, rs;  // level storage (levels >= 1 for x and b)
    std::vector<T*> xk, bk;

    Solver(int gx, int gy, int gz, bool d) : sys(levels(gx, gy, gz)), dir(d) {
        const size_t L = sys.size();
        xs.resize(L); bs.resize(L); rs.resize(L);
        xk.assign(L, nullptr); bk.assign(L, nullptr);
        for (size_t i = 0; i < L; ++i) {
            rs[i].assign(sys[i].n(), T(0));
            if (i > 0) {
                xs[i].assign(sys[i].n(), T(0));
                bs[i].assign(sys[i].n(), T(0));
                xk[i] = xs[i].data();
                bk[i] = bs[i].data();
            }
        }
    }
    int max_level() const { return (int)sys.size(); }
    void zero_from(int level) {
        for (int i = level + 1; i < max_level(); ++i) {
            std::fill(xs[i].begin(), xs[i].end(), T(0));
            std::fill(bs[i].begin(), bs[i].end(), T(0));
            std::fill(rs[i].begin(), rs[i].end(), T(0));
        }
    }
    // m_ExactSolve (S/Multigrid3D.h:1187-1191)
    void exact() { rbgs(sys[max_level() - 1], dir, bk[max_level() - 1], xk[max_level() - 1], 100, 1.0); }

    // m_Vcycle (S/Multigrid3D.h:1328-1392)
    void vcycle(T* x, const T* b, int level) {
        xk[level] = x;
        bk[level] = const_cast<T*>(b);
        zero_from(level);
        std::fill(rs[level].begin(), rs[level].end(), T(0));
        for (int i = level; i < max_level() - 1; ++i) {
            rbgs(sys[i], dir, bk[i], xk[i], 5, 1.0);
            residual(sys[i], dir, rs[i].data(), bk[i], xk[i]);
            restrict_(sys[i + 1], sys[i], rs[i].data(), bk[i + 1]);
        }
        exact();
        for (int i = max_level() - 2; i >= level; --i) {
            prolongate(sys[i], sys[i + 1], xk[i + 1], xk[i]);
            rbgs(sys[i], dir, bk[i], xk[i], 5, 1.0);
        }
        zero_from(level);
    }

    // m_FullMultiGrid (S/Multigrid3D.h:1394-1436)
    void fmg(T* x, const T* b) {
        xk[0] = x;
        bk[0] = const_cast<T*>(b);
        zero_from(0);
        std::fill(rs[0].begin(), rs[0].end(), T(0));
        for (int i = 0; i < max_level() - 1; ++i) {
            residual(sys[i], dir, rs[i].data(), bk[i], xk[i]);
            restrict_(sys[i + 1], sys[i], rs[i].data(), bk[i + 1]);
        }
        exact();
        for (int i = max_level() - 2; i >= 0; --i) {
            prolongate(sys[i], sys[i + 1], xk[i + 1], xk[i]);
            vcycle(xk[i], bk[i], i);
        }
        vcycle(xk[0], b, 0);
    }
};

template <class T> void vcycle_run(int gx, int gy, int gz, int dir, T* x, const T* b, int cycles, int fmg) {
    Solver<T> S(gx, gy, gz, dir != 0);
    if (fmg) S.fmg(x, b);
    for (int c = 0; c < cycles; ++c) S.vcycle(x, b, 0);
}

}  // namespace mg3d

extern "C" {

// Level extents, 3 per level; returns the level count (dims may be NULL).
int or_mg3d_levels(int gx, int gy, int gz, int* dims) {
    auto s = mg3d::levels(gx, gy, gz);
    if (dims)
        for (size_t l = 0; l < s.size(); ++l) {
            dims[3 * l] = s[l].x; dims[3 * l + 1] = s[l].y; dims[3 * l + 2] = s[l].z;
        }
    return (int)s.size();
}
// type 0 = float, 1 = double (MGSolverf / MGSolverd)
void or_mg3d_rbgs(int nx, int ny, int nz, int type, int dir, const void* b, void* x, int iters, double omega) {
    const mg3d::Dims g{nx, ny, nz};
    if (type) mg3d::rbgs(g, dir != 0, (const double*)b, (double*)x, iters, omega);
    else mg3d::rbgs(g, dir != 0, (const float*)b, (float*)x, iters, omega);
}
void or_mg3d_residual(int nx, int ny, int nz, int type, int dir, void* res, const void* b, const void* x) {
    const mg3d::Dims g{nx, ny, nz};
    if (type) mg3d::residual(g, dir != 0, (double*)res, (const double*)b, (const double*)x);
    else mg3d::residual(g, dir != 0, (float*)res, (const float*)b, (const float*)x);
}
void or_mg3d_restrict(int nx, int ny, int nz, int type, const void* src, void* dst) {
    const mg3d::Dims fg{nx, ny, nz};
    const mg3d::Dims cg{std::max(nx / 2, 1), std::max(ny / 2, 1), std::max(nz / 2, 1)};
    if (type) mg3d::restrict_(cg, fg, (const double*)src, (double*)dst);
    else mg3d::restrict_(cg, fg, (const float*)src, (float*)dst);
}
void or_mg3d_prolongate(int nx, int ny, int nz, int type, const void* src, void* dst) {
    const mg3d::Dims fg{nx, ny, nz};
    const mg3d::Dims cg{std::max(nx / 2, 1), std::max(ny / 2, 1), std::max(nz / 2, 1)};
    if (type) mg3d::prolongate(fg, cg, (const double*)src, (double*)dst);
    else mg3d::prolongate(fg, cg, (const float*)src, (float*)dst);
}
// fmg != 0: one m_FullMultiGrid first; then `cycles` m_Vcycle calls, one solver object.
void or_mg3d_solve(int nx, int ny, int nz, int type, int dir, void* x, const void* b, int cycles, int fmg) {
    if (type) mg3d::vcycle_run(nx, ny, nz, dir, (double*)x, (const double*)b, cycles, fmg);
    else mg3d::vcycle_run(nx, ny, nz, dir, (float*)x, (const float*)b, cycles, fmg);
}

}  // extern "C"
