/*
 * vflip_ivock.hpp — C++ drop-in facade of the reference's solver classes over the
 * C ABI of vflip.h (libvflip_ivock.so, linked with libvflip.so).
 *
 * Code written against the reference's `ivock` interface compiles against this
 * header in its place: the class names, method names and signatures, public data
 * members and the value types they use are those of
 *   SmokeSolver3D          S/Smoke_solver3D.h:25-372
 *   SubDomainSolver        S/sub_domain.h:23-290
 *   CoarseFineSmokeSolver  S/coarse_fine_solver.h:9-34
 *   AMGPCGSolve            S/AlgebraicMultigrid.h:189-197
 * and the types Vec<N,T> (S/vec.h), Buffer3D<T> (S/fluid_buffer3D.h: 8^3-brick
 * storage, element ((K*bx*by + J*bx + I) << 9) + (kk << 6) + (jj << 3) + ii),
 * particle (S/fluid_particle.h:14-65) and SparseMatrix<T> (S/sparse_matrix.h).
 * Every body here is this library's own: the fields live in host Buffer3D bricks
 * as in the reference, and each method uploads what it reads, runs the gfx950
 * kernels through vflip.h and downloads what it writes. That keeps the host
 * members authoritative between calls, as the reference code expects; callers
 * that keep fields on the device use vflip.h directly.
 *
 * Not provided: the reference's OpenGL output, its grid viscosity and
 * extrapolation (not on the hot path) and the coarse/fine curl stretching that
 * refers to members SmokeSolver3D does not have.
 */
#ifndef VFLIP_IVOCK_HPP
#define VFLIP_IVOCK_HPP

#include <cmath>
#include <algorithm>
#include <cstddef>
#include <vector>

#include "vflip.h"

namespace ivock {

typedef unsigned int uint;

// ------------------------------------------------------------------- Vec
template <unsigned N, class T> struct Vec {
    T v[N];
    Vec() {}
    Vec(T a) { for (unsigned i = 0; i < N; ++i) v[i] = a; }
    Vec(T a, T b) { v[0] = a; v[1] = b; }
    Vec(T a, T b, T c) { v[0] = a; v[1] = b; v[2] = c; }
    Vec(T a, T b, T c, T d) { v[0] = a; v[1] = b; v[2] = c; v[3] = d; }
    T& operator[](int i) { return v[i]; }
    const T& operator[](int i) const { return v[i]; }
    Vec operator+(const Vec& o) const { Vec r; for (unsigned i = 0; i < N; ++i) r.v[i] = v[i] + o.v[i]; return r; }
    Vec operator-(const Vec& o) const { Vec r; for (unsigned i = 0; i < N; ++i) r.v[i] = v[i] - o.v[i]; return r; }
    Vec operator-() const { Vec r; for (unsigned i = 0; i < N; ++i) r.v[i] = -v[i]; return r; }
    Vec operator*(T a) const { Vec r; for (unsigned i = 0; i < N; ++i) r.v[i] = v[i] * a; return r; }
    Vec operator/(T a) const { Vec r; for (unsigned i = 0; i < N; ++i) r.v[i] = v[i] / a; return r; }
    Vec& operator+=(const Vec& o) { for (unsigned i = 0; i < N; ++i) v[i] += o.v[i]; return *this; }
    Vec& operator-=(const Vec& o) { for (unsigned i = 0; i < N; ++i) v[i] -= o.v[i]; return *this; }
    Vec& operator*=(T a) { for (unsigned i = 0; i < N; ++i) v[i] *= a; return *this; }
};
template <unsigned N, class T> inline Vec<N, T> operator*(T a, const Vec<N, T>& x) { return x * a; }
typedef Vec<3, float> Vec3f;
typedef Vec<4, float> Vec4f;
typedef Vec<3, int> Vec3i;

// ------------------------------------------------------------- Buffer3D
// Host 8^3-brick field with the grid spacing and cell offsets of its samples.
template <class T> class Buffer3D {
public:
    uint _nx = 0, _ny = 0, _nz = 0, _blockx = 0, _blocky = 0, _blockz = 0;
    double _hx = 1.0, _ox = 0.0, _oy = 0.0, _oz = 0.0;
    size_t _physical_n = 0;

    void init(uint nx, uint ny, uint nz, double h, double ox, double oy, double oz) {
        _nx = nx; _ny = ny; _nz = nz;
        _blockx = (nx + 7) / 8; _blocky = (ny + 7) / 8; _blockz = (nz + 7) / 8;
        _hx = h; _ox = ox; _oy = oy; _oz = oz;
        _physical_n = (size_t)_blockx * _blocky * _blockz * 512;
        _data.assign(_physical_n, T(0));
    }
    void init(uint nx, uint ny, uint nz) { init(nx, ny, nz, 1.0, 0.0, 0.0, 0.0); }
    size_t index(int i, int j, int k) const {
        const size_t brick = ((size_t)(k >> 3) * _blocky + (size_t)(j >> 3)) * _blockx + (size_t)(i >> 3);
        return (brick << 9) + ((size_t)(k & 7) << 6) + ((size_t)(j & 7) << 3) + (size_t)(i & 7);
    }
    T& operator()(int i, int j, int k) { return _data[index(i, j, k)]; }
    const T& operator()(int i, int j, int k) const { return _data[index(i, j, k)]; }
    T at(int i, int j, int k) const {
        i = i < 0 ? 0 : (i >= (int)_nx ? (int)_nx - 1 : i);
        j = j < 0 ? 0 : (j >= (int)_ny ? (int)_ny - 1 : j);
        k = k < 0 ? 0 : (k >= (int)_nz ? (int)_nz - 1 : k);
        return (*this)(i, j, k);
    }
    void setZero() { std::fill(_data.begin(), _data.end(), T(0)); }
    void copy(const Buffer3D& b) { _data = b._data; }
    void free() { std::vector<T>().swap(_data); }
    T* data() { return _data.data(); }
    const T* data() const { return _data.data(); }

private:
    std::vector<T> _data;
};
typedef Buffer3D<float> buffer3Df;
typedef Buffer3D<char> buffer3Dc;

// ------------------------------------------------------------- particle
struct particle {
    Vec3f pos, vel;
    float dens = 0, temp = 0;
    int hash_idx = 0, sub_hash_idx = 0;
    particle() : pos(0.f), vel(0.f) {}
    particle(float px, float py, float pz, float vx, float vy, float vz, float rho, float heat)
        : pos(px, py, pz), vel(vx, vy, vz), dens(rho), temp(heat) {}
};

// --------------------------------------------------------- SparseMatrix
// Dynamic row lists, column indices ascending per row.
template <class T> struct SparseMatrix {
    unsigned n = 0;
    std::vector<std::vector<unsigned>> index;
    std::vector<std::vector<T>> value;
    explicit SparseMatrix(unsigned n_ = 0) { resize(n_); }
    void resize(unsigned n_) { n = n_; index.resize(n); value.resize(n); }
    void zero() { for (unsigned r = 0; r < n; ++r) { index[r].clear(); value[r].clear(); } }
    void clear() { n = 0; index.clear(); value.clear(); }
    T operator()(unsigned r, unsigned c) const {
        for (size_t q = 0; q < index[r].size(); ++q) if (index[r][q] == c) return value[r][q];
        return T(0);
    }
    void set_element(unsigned r, unsigned c, T v) { slot(r, c) = v; }
    void add_to_element(unsigned r, unsigned c, T v) { slot(r, c) += v; }

private:
    T& slot(unsigned r, unsigned c) {
        std::vector<unsigned>& ix = index[r];
        size_t q = 0;
        while (q < ix.size() && ix[q] < c) ++q;
        if (q == ix.size() || ix[q] != c) {
            ix.insert(ix.begin() + q, c);
            value[r].insert(value[r].begin() + q, T(0));
        }
        return value[r][q];
    }
};
typedef SparseMatrix<double> SparseMatrixd;

// ------------------------------------------------------------ AMGPCGSolve
// The drop-in of S/AlgebraicMultigrid.h:189-197: the 7-point rows of A are read
// as diag + couplings to +x/+y/+z and solved by vflip_amgpcg_solve_host (mixed
// precision by default: FP64 iterate and a stop test on the FP64 true residual).
template <class T>
bool AMGPCGSolve(const SparseMatrix<T>& A, const std::vector<T>& rhs, std::vector<T>& result,
                 T tol, int maxit, T& residual_out, int& iterations_out, int ni, int nj, int nk) {
    const size_t n = (size_t)ni * nj * nk;
    std::vector<double> d(n, 0), ax(n, 0), ay(n, 0), az(n, 0), b(rhs.begin(), rhs.end()), x(n, 0);
    for (unsigned r = 0; r < A.n; ++r)
        for (size_t q = 0; q < A.index[r].size(); ++q) {
            const unsigned c = A.index[r][q];
            const double v = (double)A.value[r][q];
            if (c == r) d[r] = v;
            else if (c == r + 1) ax[r] = v;
            else if (c == r + (unsigned)ni) ay[r] = v;
            else if (c == r + (unsigned)(ni * nj)) az[r] = v;
        }
    double res = 0;
    int its = 0;
    const int rc = vflip_amgpcg_solve_host(ni, nj, nk, d.data(), ax.data(), ay.data(), az.data(), b.data(),
                                           x.data(), (double)tol, maxit, &res, &its, nullptr);
    result.assign(x.begin(), x.end());
    residual_out = (T)res;  // absolute |r|_inf, as the reference
    iterations_out = its;
    return rc == VFLIP_OK;  // 1 = not converged, <0 = error (vflip_last_error())
}

// ---------------------------------------------------- MultiGridSolver3D
// The geometric multigrid of S/Multigrid3D.h:12-184 (uniform_grid_descriptor_3D,
// MultiGrid<T>, MultiGridSolver3D<T>, MGSolverf / MGSolverd) over vflip_mg3d_*:
// the same levels, members and methods; each method uploads the Buffer3D fields it
// reads, runs the gfx950 kernels and downloads what it writes (bit-identical with
// the reference's arithmetic in T). m_Vcycle / m_FullMultiGrid take tol and residual
// and leave them unused, as the reference does. The open-boundary far-field members
// (m_applyOpenBoundaryCondition, m_FastSummation, m_BarnesHutSummation, m_Toplevel,
// m_bottomlevel, m_cumulate_poles, m_MonopleToMonopole) are not provided.
struct uniform_grid_descriptor_3D {
    int gridx, gridy, gridz, system_size;
    uniform_grid_descriptor_3D() : gridx(0), gridy(0), gridz(0), system_size(0) {}
    uniform_grid_descriptor_3D(int x, int y, int z) : gridx(x), gridy(y), gridz(z), system_size(x * y * z) {}
    void set_grid_information(int x, int y, int z) {
        gridx = x; gridy = y; gridz = z;
        system_size = x * y * z;
    }
};

template <class T> class MultiGrid {
public:
    MultiGrid() : m_bInitialized(false), m_max_level(0) {}
    virtual ~MultiGrid() { m_FinalMemoryEachLevel(); }
    void m_InitialSystem(int gridx, int gridy, int gridz);
    void m_SetGridInformation(int x_, int y_, int z_) { m_system_descriptor.set_grid_information(x_, y_, z_); }
    void m_AssignMemoryEachLevel();
    void m_FinalMemoryEachLevel();
    void m_ComputeLevels();
    bool m_bInitialized;
    uniform_grid_descriptor_3D m_system_descriptor;
    int m_max_level;
    std::vector<Buffer3D<T>*> xk;
    std::vector<Buffer3D<T>*> xk_new;
    std::vector<Buffer3D<T>*> bk;
    std::vector<Buffer3D<T>*> rk;
    std::vector<uniform_grid_descriptor_3D> systemk;

protected:
    std::vector<Buffer3D<T>*> owned_;  // level buffers this object allocated
};

struct MGDevice;  // device copies and vflip_mg3d handles (facade implementation)

template <class T> class MultiGridSolver3D : public MultiGrid<T> {
public:
    MultiGridSolver3D() : _b_Dirichlet(false) {}
    ~MultiGridSolver3D();
    bool _b_Dirichlet;
    void m_Vcycle(Buffer3D<T>* x, Buffer3D<T>* b, T tol, T& residual, int level);
    void m_FullMultiGrid(Buffer3D<T>* x, Buffer3D<T>* b, T tol, T& residual);
    void m_Red_Black_Gauss(uniform_grid_descriptor_3D& system, Buffer3D<T>& b, Buffer3D<T>& x, Buffer3D<T>& x_new,
                           int iter_time, double omega);
    void m_Restrict(uniform_grid_descriptor_3D& next_level, uniform_grid_descriptor_3D& curr_level,
                    Buffer3D<T>& src_level, Buffer3D<T>& dst_level);
    void m_ExactSolve();
    void m_Prolongate(uniform_grid_descriptor_3D& next_level, uniform_grid_descriptor_3D& curr_level,
                      Buffer3D<T>& src_level, Buffer3D<T>& dst_level);
    void m_ComputeResidual(uniform_grid_descriptor_3D& system, Buffer3D<T>& res, Buffer3D<T>& b, Buffer3D<T>& x);

private:
    MGDevice* dev_ = nullptr;
    MGDevice& dev();
};
typedef MultiGridSolver3D<float> MGSolverf;
typedef MultiGridSolver3D<double> MGSolverd;

struct DeviceState;  // device fields and vflip handles of one solver (facade implementation)

// --------------------------------------------------------- SubDomainSolver
class SubDomainSolver {
public:
    SubDomainSolver();
    ~SubDomainSolver();
    Vec3f bmin, bmax;
    uint _nx = 0, _ny = 0, _nz = 0, _n = 0;
    float _hx = 0, _hy = 0, _hz = 0, _lx = 0, _ly = 0, _lz = 0;
    buffer3Df _un, _vn, _wn, _utemp, _vtemp, _wtemp, _unp1, _vnp1, _wnp1;
    buffer3Df _u_weights, _v_weights, _w_weights;
    buffer3Df _u_solid, _v_solid, _w_solid;
    buffer3Df _nodal_solid_phi;
    std::vector<double> rhs, pressure;
    std::vector<particle> FLIP;
    vflip_solve_stats last_stats{};
    unsigned seed = 7;  // reseed draws (counter hash in place of rand()); facade state

    void init(Vec3f& _bmin, Vec3f& _bmax, float dx);
    bool is_in(Vec3f& pos);
    bool is_in2(Vec3f& pos);
    bool is_halo(Vec3f& pos, float dx);
    void set_boundary_phi(float (*phi)(const Vec3f&));
    void set_boundary_vel(Vec3f (*vel)(const Vec3f&));
    void compute_face_weights();
    void pcg_projection(float dt);
    void advect_field(float dt, buffer3Df& field, buffer3Df& field_new);
    void getDelta(buffer3Df& field_old, buffer3Df& field_new, buffer3Df& fieldtemp);
    void compute_delta_u();
    void hash_FLIP();
    void particle_to_grid(buffer3Df& field, buffer3Df& coef, int component);
    void particle_to_grid();
    void particle_from_grid();
    void reorder_FLIP();
    void Finalize();

    DeviceState* dev = nullptr;  // facade state (not part of the reference interface)
};

// ----------------------------------------------------------- SmokeSolver3D
class SmokeSolver3D {
public:
    SmokeSolver3D();
    ~SmokeSolver3D();
    SubDomainSolver sub_domain;
    Vec3f emitter_pos;
    float emitter_r = 0;
    uint emitter_n = 0;
    std::vector<Vec4f> tracers;
    Vec3f origin;
    uint _nx = 0, _ny = 0, _nz = 0, _n = 0;
    float _hx = 0, _hy = 0, _hz = 0, _lx = 0, _ly = 0, _lz = 0;
    float _temp_decay = 0, _dens_decay = 0, _alpha = 0, _beta = 0;
    float _smoke_heat = 0, _smoke_dens = 0, _smoke_fuel = 0, _smoke_radius = 0, _U_in = 0;
    Vec3f _smoke_center;
    buffer3Df _un, _vn, _wn, _utemp, _vtemp, _wtemp, _unp1, _vnp1, _wnp1;
    buffer3Df _u_weights, _v_weights, _w_weights;
    buffer3Df _u_solid, _v_solid, _w_solid;
    buffer3Df _nodal_solid_phi;
    buffer3Dc _b_desc, _h_desc;
    float _cfl = 0, nu = 0;
    buffer3Df _rho, _Temperature, _temp;
    std::vector<double> rhs, pressure;
    std::vector<particle> FLIP;
    vflip_solve_stats last_stats{};

    void init(uint nx, uint ny, uint nz, float L, float ratio);
    void initSubdomain(Vec3f& submin, Vec3f& submax, float h);
    void setEmitter(Vec3f& pos, float r, uint n) { emitter_pos = pos; emitter_r = r; emitter_n = n; }
    void setSmoke(double temp_decay, double alpha, double beta, double smoke_heat, double smoke_dens) {
        _temp_decay = (float)temp_decay; _alpha = (float)alpha; _beta = (float)beta;
        _smoke_heat = (float)smoke_heat; _smoke_dens = (float)smoke_dens;
    }
    void setSmokeDens(float v) { _smoke_dens = v; }
    void setSmokeHeat(float v) { _smoke_heat = v; }
    void setSmokeHeatDecay(float v) { _temp_decay = v; }
    void setSmokeDensDecay(float v) { _dens_decay = v; }
    void setSmokeRadius(float v) { _smoke_radius = v; }
    void setSmokeCenter(Vec3f& v) { _smoke_center = v; }
    void setDensBuoy(float v) { _alpha = v; }
    void setTempBuoy(float v) { _beta = v; }
    void set_viscosity(float v) { nu = v; }
    void set_U_in(float v) { _U_in = v; }
    void set_boundary(buffer3Dc& b_desc) { _b_desc.copy(b_desc); }
    void set_heat(buffer3Dc& h_desc) { _h_desc.copy(h_desc); }
    void set_boundary_phi(float (*phi)(const Vec3f&));
    void set_boundary_vel(Vec3f (*vel)(const Vec3f&));
    void compute_face_weights();
    void pcg_projection(float dt);
    void advect_field(float dt, buffer3Df& field, buffer3Df& field_new);
    void advect_field_cubic_clamp(float dt, buffer3Df& field, buffer3Df& field_new);
    void getDelta(buffer3Df& field_old, buffer3Df& field_new, buffer3Df& fieldtemp);
    void compute_delta_u();
    void hash_FLIP();
    void particle_to_grid(buffer3Df& field, buffer3Df& coef, int component);
    void particle_from_grid();
    void advect_FLIP(float dt);
    void reorder_FLIP();
    void seed_ghost();
    void diffuse(float dt);
    void emit_tracers();
    void advect_tracers(float dt);
    void write_tracers(const char* file_path, int frame);
    void getCFL(float dt);
    void time_step(float dt, int adv_type, int frame);
    void Finalize();

    // facade options (not part of the reference interface): P2G summation order
    // (true: the reference's order, bit-exact; false: owner-computes tiles) and the
    // projection's precision (VFLIP_PREC_MIXED by default: FP64 stop test, as the
    // reference's FP64 solve)
    bool exact_transfers = true;
    int precision = VFLIP_PREC_MIXED;
    unsigned seed = 1;
    DeviceState* dev = nullptr;
};

// --------------------------------------------------- CoarseFineSmokeSolver
class CoarseFineSmokeSolver {
public:
    CoarseFineSmokeSolver();
    ~CoarseFineSmokeSolver();
    void Finalize();
    uint _nx = 0, _ny = 0, _nz = 0, _N = 8;
    float _h = 0;
    SmokeSolver3D fine_solver;
    SmokeSolver3D coarse_solver;
    void initialize(uint nx, uint ny, uint nz, float L);
    void FineProjection();
    void SetFaceBoundary(buffer3Df& fine, buffer3Df& coarse, int offx, int offy, int offz);
    void FineToCoarseOperator(buffer3Df& fine, buffer3Df& coarse, int offx, int offy, int offz);
    void FineToCoarse();
    void CoarseToFineOperator(buffer3Df& fine, buffer3Df& coarse);
    void ComputeSmallDifference(buffer3Df& fine, buffer3Df& coarse, buffer3Df& diff);
    void LargeScaleAdvection(float dt);
};

}  // namespace ivock

#endif  // VFLIP_IVOCK_HPP
