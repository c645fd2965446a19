// Method bodies of the C++ drop-in facade (include/vflip_ivock.hpp) over the C ABI of
// include/vflip.h. Host Buffer3D bricks stay the authoritative copies between calls:
// each method uploads the fields it reads (bricks -> linear on the GPU), runs the
// library's kernels and downloads what it writes. Reference method each body stands
// in for: S/Smoke_solver3D.{h,cpp}, S/sub_domain.{h,cpp}, S/coarse_fine_solver.cpp,
// cited per method.
#include "vflip_ivock.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace ivock {

namespace {

void fail(const char* what, int rc) {
    char msg[512];
    std::snprintf(msg, sizeof msg, "vflip facade: %s failed (%d): %s", what, rc, vflip_last_error());
    throw std::runtime_error(msg);
}
void ck(int rc, const char* what) {
    if (rc < 0) fail(what, rc);
}
void hk(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        char msg[512];
        std::snprintf(msg, sizeof msg, "vflip facade: %s: %s", what, hipGetErrorString(e));
        throw std::runtime_error(msg);
    }
}

// Device buffer that grows on demand.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void* get(size_t b) {
        if (b > bytes) {
            if (p) (void)hipFree(p);
            p = nullptr;
            hk(hipMalloc(&p, std::max<size_t>(b, 256)), "hipMalloc");
            bytes = std::max<size_t>(b, 256);
        }
        return p;
    }
    template <class T> T* as(size_t n) { return (T*)get(n * sizeof(T)); }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

// Device copies and handles of one solver's grid (coarse or subdomain).
struct DeviceState {
    int ni = 0, nj = 0, nk = 0;
    float h = 0;
    vflip_solver* S = nullptr;
    vflip_particles* P = nullptr;
    DBuf u, v, w, uw, vw, ww, us, vs, ws, phi, rhs, p, bdesc, bricks, tmp0, tmp1, tmp2;
    DBuf pos, vel, pos2, vel2, key, start, order, tr;
    int np_binned = -1;  // particles of the last binning
    size_t nu() const { return (size_t)(ni + 1) * nj * nk; }
    size_t nv() const { return (size_t)ni * (nj + 1) * nk; }
    size_t nw() const { return (size_t)ni * nj * (nk + 1); }
    size_t nc() const { return (size_t)ni * nj * nk; }
    size_t nn() const { return (size_t)(ni + 1) * (nj + 1) * (nk + 1); }
    ~DeviceState() {
        if (S) vflip_solver_destroy(S);
        if (P) vflip_particles_destroy(P);
    }
};

namespace {

DeviceState* make_state(int ni, int nj, int nk, float h, int precision, double tol, int maxit) {
    DeviceState* D = new DeviceState();
    D->ni = ni; D->nj = nj; D->nk = nk; D->h = h;
    vflip_solve_opts o;
    vflip_default_opts(&o);
    o.precision = precision;
    o.tolerance = tol;
    o.max_iterations = maxit;
    ck(vflip_solver_create(ni, nj, nk, &o, &D->S), "solver_create");
    ck(vflip_particles_create(ni, nj, nk, h, &D->P), "particles_create");
    return D;
}

// Buffer3D bricks -> linear device array (n = nx ny nz floats at dst).
void upload(DeviceState* D, const buffer3Df& b, float* dst) {
    float* br = D->bricks.as<float>(b._physical_n);
    hk(hipMemcpy(br, b.data(), b._physical_n * sizeof(float), hipMemcpyHostToDevice), "upload");
    ck(vflip_from_bricks(br, (int)b._nx, (int)b._ny, (int)b._nz, dst, nullptr), "from_bricks");
}
// linear device array -> Buffer3D bricks
void download(DeviceState* D, const float* src, buffer3Df& b) {
    float* br = D->bricks.as<float>(b._physical_n);
    ck(vflip_to_bricks(src, (int)b._nx, (int)b._ny, (int)b._nz, br, nullptr), "to_bricks");
    hk(hipMemcpy(b.data(), br, b._physical_n * sizeof(float), hipMemcpyDeviceToHost), "download");
}
void upload_bdesc(DeviceState* D, const buffer3Dc& b) {
    std::vector<int8_t> lin((size_t)b._nx * b._ny * b._nz);
    for (uint k = 0; k < b._nz; ++k)
        for (uint j = 0; j < b._ny; ++j)
            for (uint i = 0; i < b._nx; ++i) lin[(size_t)i + (size_t)b._nx * (j + (size_t)b._ny * k)] = (int8_t)b(i, j, k);
    hk(hipMemcpy(D->bdesc.get(lin.size()), lin.data(), lin.size(), hipMemcpyHostToDevice), "upload bdesc");
}
// particle positions / velocities <-> xyz-interleaved device arrays
void upload_particles(DeviceState* D, const std::vector<particle>& F, bool vel) {
    const size_t n = F.size();
    std::vector<float> h(3 * std::max<size_t>(n, 1));
    for (size_t q = 0; q < n; ++q)
        for (int a = 0; a < 3; ++a) h[3 * q + a] = F[q].pos[a];
    hk(hipMemcpy(D->pos.as<float>(h.size()), h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload pos");
    if (vel) {
        for (size_t q = 0; q < n; ++q)
            for (int a = 0; a < 3; ++a) h[3 * q + a] = F[q].vel[a];
        hk(hipMemcpy(D->vel.as<float>(h.size()), h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload vel");
    }
}
void download_particles(const float* dpos, const float* dvel, size_t n, std::vector<particle>& F) {
    std::vector<float> a(3 * std::max<size_t>(n, 1)), b(3 * std::max<size_t>(n, 1));
    if (dpos) hk(hipMemcpy(a.data(), dpos, 3 * n * 4, hipMemcpyDeviceToHost), "download pos");
    if (dvel) hk(hipMemcpy(b.data(), dvel, 3 * n * 4, hipMemcpyDeviceToHost), "download vel");
    F.resize(n);
    for (size_t q = 0; q < n; ++q)
        for (int c = 0; c < 3; ++c) {
            if (dpos) F[q].pos[c] = a[3 * q + c];
            if (dvel) F[q].vel[c] = b[3 * q + c];
        }
}
void bin(DeviceState* D, int np, int clamp, const Vec3f& bmin) {
    ck(vflip_hash(D->P, np, np ? D->pos.as<float>(3 * np) : nullptr, clamp, bmin[0], bmin[1], bmin[2],
                  D->key.as<int>(std::max(np, 1)), D->start.as<int>(D->nc() + 1), D->order.as<int>(std::max(np, 1)),
                  nullptr),
       "hash");
    D->np_binned = np;
}
// Component c's MAC face grid dims.
void face_dims(const DeviceState* D, int c, int& x, int& y, int& z) {
    x = D->ni + (c == 0); y = D->nj + (c == 1); z = D->nk + (c == 2);
}

// phi / vel callbacks of the reference (Vec3f by reference) behind the C callbacks of
// vflip_set_boundary_phi / vflip_set_boundary_vel; the subdomain adds its bmin to the
// grid point as the reference does (pos = Vec3f(i,j,k)*h - o*h + bmin).
struct PhiCall {
    float (*fn)(const Vec3f&);
    Vec3f shift;
    bool shifted;
};
struct VelCall {
    Vec3f (*fn)(const Vec3f&);
    Vec3f shift;
    bool shifted;
};
float phi_tramp(const float* pos, void* user) {
    const PhiCall* c = (const PhiCall*)user;
    Vec3f p(pos[0], pos[1], pos[2]);
    if (c->shifted) p = p + c->shift;
    return c->fn(p);
}
void vel_tramp(const float* pos, float* out, void* user) {
    const VelCall* c = (const VelCall*)user;
    Vec3f p(pos[0], pos[1], pos[2]);
    if (c->shifted) p = p + c->shift;
    const Vec3f v = c->fn(p);
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
}

// The MAC fields of a solver's grid, uploaded: velocity, solid velocity, nodal phi.
template <class Solver> void upload_flow(DeviceState* D, Solver& s) {
    upload(D, s._un, D->u.as<float>(D->nu()));
    upload(D, s._vn, D->v.as<float>(D->nv()));
    upload(D, s._wn, D->w.as<float>(D->nw()));
    upload(D, s._u_solid, D->us.as<float>(D->nu()));
    upload(D, s._v_solid, D->vs.as<float>(D->nv()));
    upload(D, s._w_solid, D->ws.as<float>(D->nw()));
    upload(D, s._nodal_solid_phi, D->phi.as<float>(D->nn()));
}
template <class Solver> void download_velocity(DeviceState* D, Solver& s) {
    download(D, (float*)D->u.p, s._un);
    download(D, (float*)D->v.p, s._vn);
    download(D, (float*)D->w.p, s._wn);
}
template <class Solver> void face_weights(DeviceState* D, Solver& s) {
    upload(D, s._nodal_solid_phi, D->phi.as<float>(D->nn()));
    ck(vflip_face_weights(D->ni, D->nj, D->nk, (float*)D->phi.p, D->uw.as<float>(D->nu()), D->vw.as<float>(D->nv()),
                          D->ww.as<float>(D->nw()), nullptr),
       "face_weights");
    download(D, (float*)D->uw.p, s._u_weights);
    download(D, (float*)D->vw.p, s._v_weights);
    download(D, (float*)D->ww.p, s._w_weights);
}
// The MAC buffers of a grid with the reference's offsets (S/Smoke_solver3D.h:299-330).
template <class Solver> void init_buffers(Solver& s) {
    const uint nx = s._nx, ny = s._ny, nz = s._nz;
    const double h = s._hx;
    for (buffer3Df* b : {&s._un, &s._utemp, &s._unp1, &s._u_solid}) b->init(nx + 1, ny, nz, h, 0.5, 0, 0);
    for (buffer3Df* b : {&s._vn, &s._vtemp, &s._vnp1, &s._v_solid}) b->init(nx, ny + 1, nz, h, 0.0, 0.5, 0.0);
    for (buffer3Df* b : {&s._wn, &s._wtemp, &s._wnp1, &s._w_solid}) b->init(nx, ny, nz + 1, h, 0.0, 0.0, 0.5);
    s._u_weights.init(nx + 1, ny, nz);
    s._v_weights.init(nx, ny + 1, nz);
    s._w_weights.init(nx, ny, nz + 1);
    s._nodal_solid_phi.init(nx + 1, ny + 1, nz + 1, h, 0.5, 0.5, 0.5);
}
// getDelta of S/Smoke_solver3D.cpp:1298-1318: temp = new - old, element by element
// (the brick layout is the same for the three buffers).
void delta(DeviceState* D, buffer3Df& old_, buffer3Df& new_, buffer3Df& temp) {
    const size_t n = temp._physical_n;
    float* a = D->tmp0.as<float>(n);
    float* b = D->tmp1.as<float>(n);
    float* c = D->tmp2.as<float>(n);
    hk(hipMemcpy(a, old_.data(), n * 4, hipMemcpyHostToDevice), "upload");
    hk(hipMemcpy(b, new_.data(), n * 4, hipMemcpyHostToDevice), "upload");
    ck(vflip_get_delta(n, b, a, c, nullptr), "get_delta");
    hk(hipMemcpy(temp.data(), c, n * 4, hipMemcpyDeviceToHost), "download");
}
vflip_coarse_view coarse_view(DeviceState* D) {
    return vflip_coarse_view{D->ni, D->nj, D->nk, D->h, (float*)D->u.p, (float*)D->v.p, (float*)D->w.p,
                             (float*)D->phi.p, (float*)D->us.p, (float*)D->vs.p, (float*)D->ws.p};
}
vflip_subdomain_view sub_view(DeviceState* D, const SubDomainSolver& s) {
    vflip_subdomain_view v{D->ni, D->nj, D->nk, D->h, {s.bmin[0], s.bmin[1], s.bmin[2]},
                           {s.bmax[0], s.bmax[1], s.bmax[2]}, (float*)D->u.p, (float*)D->v.p, (float*)D->w.p,
                           (float*)D->phi.p, (float*)D->us.p, (float*)D->vs.p, (float*)D->ws.p};
    return v;
}
// Grows a selection's output until it fits (VFLIP_ERR_CAPACITY reports the size needed).
template <class F> int select_into(DeviceState* D, int cap, F call) {
    for (int attempt = 0; attempt < 2; ++attempt) {
        int n = 0;
        const int rc = call(cap, D->pos2.as<float>(3 * (size_t)std::max(cap, 1)),
                            D->vel2.as<float>(3 * (size_t)std::max(cap, 1)), &n);
        if (rc == VFLIP_ERR_CAPACITY && attempt == 0) {
            cap = n;
            continue;
        }
        ck(rc, "selection");
        return n;
    }
    return 0;
}

}  // namespace

// ======================================================== SubDomainSolver
SubDomainSolver::SubDomainSolver() {}
SubDomainSolver::~SubDomainSolver() { Finalize(); }

void SubDomainSolver::Finalize() {
    delete dev;
    dev = nullptr;
}

// S/sub_domain.h:160-176 and 218-258
void SubDomainSolver::init(Vec3f& _bmin, Vec3f& _bmax, float dx) {
    bmin = _bmin;
    bmax = _bmax;
    float L = bmax[0] - bmin[0];
    const int nx = (int)std::floor(L / dx);
    const int ny = (int)std::floor((bmax[1] - bmin[1]) / dx);
    const int nz = (int)std::floor((bmax[2] - bmin[2]) / dx);
    L = (float)nx * dx;
    bmax = bmin + Vec3f((float)nx, (float)ny, (float)nz) * dx;
    _nx = nx; _ny = ny; _nz = nz; _n = _nx * _ny * _nz;
    _hx = L / (float)_nx;
    _hz = _hy = _hx;
    _lx = L; _ly = _hy * (float)ny; _lz = _hz * (float)nz;
    init_buffers(*this);
    Finalize();
    // the subdomain solve: FP64 stop test to 1e-9, at most 1000 iterations (S/sub_domain.cpp:739)
    dev = make_state(nx, ny, nz, _hx, VFLIP_PREC_MIXED, 1e-9, 1000);
    hash_FLIP();
    reorder_FLIP();
}

bool SubDomainSolver::is_in(Vec3f& p) {
    for (int a = 0; a < 3; ++a)
        if (!(p[a] > bmin[a] + 2.0 * _hx && p[a] < bmax[a] - 2.0 * _hx)) return false;
    return true;
}
bool SubDomainSolver::is_in2(Vec3f& p) {
    for (int a = 0; a < 3; ++a)
        if (!(p[a] > bmin[a] + 5.0 * _hx && p[a] < bmax[a] - 5.0 * _hx)) return false;
    return true;
}
bool SubDomainSolver::is_halo(Vec3f& p, float dx) {
    for (int a = 0; a < 3; ++a)
        if (!(p[a] > bmin[a] - dx && p[a] < bmax[a] + dx)) return false;
    return true;
}

// S/sub_domain.cpp:478-490: nodes at Vec3f(i,j,k)*h - Vec3f(.5)*h + bmin
void SubDomainSolver::set_boundary_phi(float (*phi)(const Vec3f&)) {
    PhiCall c{phi, bmin, true};
    ck(vflip_set_boundary_phi(dev->ni, dev->nj, dev->nk, _hx, phi_tramp, &c, dev->phi.as<float>(dev->nn()), nullptr),
       "set_boundary_phi");
    download(dev, (float*)dev->phi.p, _nodal_solid_phi);
}
void SubDomainSolver::set_boundary_vel(Vec3f (*vel)(const Vec3f&)) {
    VelCall c{vel, bmin, true};
    ck(vflip_set_boundary_vel(dev->ni, dev->nj, dev->nk, _hx, vel_tramp, &c, dev->us.as<float>(dev->nu()),
                              dev->vs.as<float>(dev->nv()), dev->ws.as<float>(dev->nw()), nullptr),
       "set_boundary_vel");
    download(dev, (float*)dev->us.p, _u_solid);
    download(dev, (float*)dev->vs.p, _v_solid);
    download(dev, (float*)dev->ws.p, _w_solid);
}
void SubDomainSolver::compute_face_weights() { face_weights(dev, *this); }

// SubDomainSolver::pcg_projection, S/sub_domain.cpp:556-820 (dt in the matrix and the
// gradient, faces 1..n-1; FP64 stop test to 1e-9, 1000 iterations)
void SubDomainSolver::pcg_projection(float dt) {
    face_weights(dev, *this);
    upload_flow(dev, *this);
    ck(vflip_project(dev->S, VFLIP_VARIANT_SUB, _hx, dt, (float*)dev->u.p, (float*)dev->v.p, (float*)dev->w.p,
                     (float*)dev->uw.p, (float*)dev->vw.p, (float*)dev->ww.p, (float*)dev->us.p, (float*)dev->vs.p,
                     (float*)dev->ws.p, dev->rhs.as<float>(dev->nc()), dev->p.as<float>(dev->nc()), &last_stats),
       "project (subdomain)");
    download_velocity(dev, *this);
    std::vector<float> p(dev->nc()), r(dev->nc());
    hk(hipMemcpy(p.data(), dev->p.p, p.size() * 4, hipMemcpyDeviceToHost), "download pressure");
    hk(hipMemcpy(r.data(), dev->rhs.p, r.size() * 4, hipMemcpyDeviceToHost), "download rhs");
    pressure.assign(p.begin(), p.end());
    rhs.assign(r.begin(), r.end());
}

// SubDomainSolver::advect_field, S/sub_domain.cpp:191-235 (RK2 backtrace, trilinear)
void SubDomainSolver::advect_field(float dt, buffer3Df& field, buffer3Df& field_new) {
    if (std::fabs(field._hx - (double)_hx) > 1e-6 * _hx) fail("advect_field: field spacing differs from the grid's", -1);
    upload_flow(dev, *this);
    const size_t n = (size_t)field._nx * field._ny * field._nz;
    float* f = dev->tmp0.as<float>(n);
    float* o = dev->tmp1.as<float>(n);
    upload(dev, field, f);
    ck(vflip_advect_rk2(dev->ni, dev->nj, dev->nk, _hx, dt, 0, (float*)dev->u.p, (float*)dev->v.p, (float*)dev->w.p,
                        f, (int)field._nx, (int)field._ny, (int)field._nz, (float)field._ox, (float)field._oy,
                        (float)field._oz, o, nullptr, nullptr, nullptr),
       "advect_rk2");
    download(dev, o, field_new);
}

void SubDomainSolver::getDelta(buffer3Df& o, buffer3Df& n, buffer3Df& t) { delta(dev, o, n, t); }
void SubDomainSolver::compute_delta_u() {
    getDelta(_utemp, _un, _utemp);
    getDelta(_vtemp, _vn, _vtemp);
    getDelta(_wtemp, _wn, _wtemp);
}

// S/sub_domain.cpp:1087-1107: positions relative to bmin, cell indices clamped
void SubDomainSolver::hash_FLIP() {
    upload_particles(dev, FLIP, true);
    bin(dev, (int)FLIP.size(), 1, bmin);
}

// S/sub_domain.cpp:1204-1265
void SubDomainSolver::particle_to_grid(buffer3Df& field, buffer3Df& coef, int component) {
    if (dev->np_binned != (int)FLIP.size()) hash_FLIP();
    upload_particles(dev, FLIP, true);
    int fx, fy, fz;
    face_dims(dev, component, fx, fy, fz);
    const size_t nf = (size_t)fx * fy * fz;
    float* f = dev->tmp0.as<float>(nf);
    float* c = dev->tmp1.as<float>(nf);
    float* F[3] = {nullptr, nullptr, nullptr};
    float* C[3] = {nullptr, nullptr, nullptr};
    F[component] = f;
    C[component] = c;
    const int np = (int)FLIP.size();
    ck(vflip_p2g(dev->P, VFLIP_VARIANT_SUB, 1, np, np ? (float*)dev->pos.p : nullptr,
                 np ? (float*)dev->vel.p : nullptr, (int*)dev->start.p, (int*)dev->order.p, bmin[0], bmin[1], bmin[2],
                 F[0], C[0], F[1], C[1], F[2], C[2]),
       "p2g (subdomain)");
    download(dev, f, field);
    download(dev, c, coef);
}
// S/sub_domain.h:100-110
void SubDomainSolver::particle_to_grid() {
    _utemp.setZero();
    _vtemp.setZero();
    _wtemp.setZero();
    particle_to_grid(_un, _unp1, 0);
    particle_to_grid(_vn, _vnp1, 1);
    particle_to_grid(_wn, _wnp1, 2);
    _utemp.copy(_un);
    _vtemp.copy(_vn);
    _wtemp.copy(_wn);
}
// vel += trilinear(_utemp, _vtemp, _wtemp) at pos - bmin (the deltas of compute_delta_u)
void SubDomainSolver::particle_from_grid() {
    const int np = (int)FLIP.size();
    if (!np) return;
    upload_particles(dev, FLIP, true);
    float* du = dev->u.as<float>(dev->nu());
    float* dv = dev->v.as<float>(dev->nv());
    float* dw = dev->w.as<float>(dev->nw());
    upload(dev, _utemp, du);
    upload(dev, _vtemp, dv);
    upload(dev, _wtemp, dw);
    ck(vflip_g2p(dev->ni, dev->nj, dev->nk, _hx, np, (float*)dev->pos.p, (float*)dev->vel.p, du, dv, dw, nullptr,
                 nullptr, nullptr, 1.0f, bmin[0], bmin[1], bmin[2], nullptr),
       "g2p (subdomain)");
    download_particles(nullptr, (float*)dev->vel.p, (size_t)np, FLIP);
}
// S/sub_domain.cpp:1110-1159
void SubDomainSolver::reorder_FLIP() {
    hash_FLIP();
    upload_flow(dev, *this);
    const int np = (int)FLIP.size();
    const int cap0 = np + 8 * (int)dev->nc();
    const int n = select_into(dev, cap0, [&](int cap, float* op, float* ov, int* nout) {
        return vflip_reorder(dev->P, VFLIP_VARIANT_SUB, bmin[0], bmin[1], bmin[2], seed, np,
                             np ? (float*)dev->pos.p : nullptr, np ? (float*)dev->vel.p : nullptr, (int*)dev->start.p,
                             np ? (int*)dev->order.p : nullptr, (float*)dev->u.p, (float*)dev->v.p, (float*)dev->w.p,
                             (float*)dev->phi.p, cap, op, ov, nout);
    });
    ++seed;
    download_particles((float*)dev->pos2.p, (float*)dev->vel2.p, (size_t)n, FLIP);
    dev->np_binned = -1;
}

// ========================================================== SmokeSolver3D
SmokeSolver3D::SmokeSolver3D() {}
SmokeSolver3D::~SmokeSolver3D() { Finalize(); }

void SmokeSolver3D::Finalize() {
    delete dev;
    dev = nullptr;
}

// S/Smoke_solver3D.h:259-330
void SmokeSolver3D::init(uint nx, uint ny, uint nz, float L, float /*ratio*/) {
    origin = Vec3f(0.f);
    _nx = nx; _ny = ny; _nz = nz; _n = nx * ny * nz;
    _hx = L / (float)_nx;
    _hz = _hy = _hx;
    _lx = L; _ly = _hy * (float)ny; _lz = _hz * (float)nz;
    _temp_decay = 0;
    _alpha = _beta = 0;
    _smoke_dens = _smoke_heat = 0;
    init_buffers(*this);
    _rho.init(nx * 2, ny * 2, nz * 2, 0.5 * _hx, 0.0, 0.0, 0.0);
    _Temperature.init(nx * 2, ny * 2, nz * 2, 0.5 * _hx, 0.0, 0.0, 0.0);
    _temp.init(nx * 2, ny * 2, nz * 2, 0.5 * _hx, 0.0, 0.0, 0.0);
    _b_desc.init(nx, ny, nz);
    _h_desc.init(nx, ny, nz);
    Finalize();
    dev = make_state((int)nx, (int)ny, (int)nz, _hx, precision, 1e-6, 50);
    hash_FLIP();
    reorder_FLIP();
}

void SmokeSolver3D::initSubdomain(Vec3f& submin, Vec3f& submax, float h) { sub_domain.init(submin, submax, h); }

// S/Smoke_solver3D.cpp:851-873 (and the subdomain's nodes)
void SmokeSolver3D::set_boundary_phi(float (*phi)(const Vec3f&)) {
    PhiCall c{phi, Vec3f(0.f), false};
    ck(vflip_set_boundary_phi(dev->ni, dev->nj, dev->nk, _hx, phi_tramp, &c, dev->phi.as<float>(dev->nn()), nullptr),
       "set_boundary_phi");
    download(dev, (float*)dev->phi.p, _nodal_solid_phi);
    if (sub_domain.dev) sub_domain.set_boundary_phi(phi);
}
// S/Smoke_solver3D.cpp:793-848
void SmokeSolver3D::set_boundary_vel(Vec3f (*vel)(const Vec3f&)) {
    VelCall c{vel, Vec3f(0.f), false};
    ck(vflip_set_boundary_vel(dev->ni, dev->nj, dev->nk, _hx, vel_tramp, &c, dev->us.as<float>(dev->nu()),
                              dev->vs.as<float>(dev->nv()), dev->ws.as<float>(dev->nw()), nullptr),
       "set_boundary_vel");
    download(dev, (float*)dev->us.p, _u_solid);
    download(dev, (float*)dev->vs.p, _v_solid);
    download(dev, (float*)dev->ws.p, _w_solid);
    if (sub_domain.dev) sub_domain.set_boundary_vel(vel);
}
// S/Smoke_solver3D.cpp:874-944
void SmokeSolver3D::compute_face_weights() { face_weights(dev, *this); }

// S/Smoke_solver3D.cpp:945-1200: face weights, inflow faces, assembly, AMGPCG, gradient
void SmokeSolver3D::pcg_projection(float dt) {
    face_weights(dev, *this);
    upload_flow(dev, *this);
    upload_bdesc(dev, _b_desc);
    ck(vflip_inflow_bc(dev->ni, dev->nj, dev->nk, (int8_t*)dev->bdesc.p, _U_in, (float*)dev->u.p, (float*)dev->v.p,
                       (float*)dev->w.p, nullptr),
       "inflow_bc");
    ck(vflip_project(dev->S, VFLIP_VARIANT_COARSE, _hx, dt, (float*)dev->u.p, (float*)dev->v.p, (float*)dev->w.p,
                     (float*)dev->uw.p, (float*)dev->vw.p, (float*)dev->ww.p, (float*)dev->us.p, (float*)dev->vs.p,
                     (float*)dev->ws.p, dev->rhs.as<float>(dev->nc()), dev->p.as<float>(dev->nc()), &last_stats),
       "project");
    download_velocity(dev, *this);
    std::vector<float> p(dev->nc()), r(dev->nc());
    hk(hipMemcpy(p.data(), dev->p.p, p.size() * 4, hipMemcpyDeviceToHost), "download pressure");
    hk(hipMemcpy(r.data(), dev->rhs.p, r.size() * 4, hipMemcpyDeviceToHost), "download rhs");
    pressure.assign(p.begin(), p.end());
    rhs.assign(r.begin(), r.end());
}

namespace {
void advect_mode(DeviceState* D, SmokeSolver3D& s, float dt, int mode, buffer3Df& field, buffer3Df& out) {
    if (std::fabs(field._hx - (double)s._hx) > 1e-6 * s._hx)
        fail("advect_field: field spacing differs from the velocity grid's", -1);
    upload_flow(D, s);
    const size_t n = (size_t)field._nx * field._ny * field._nz;
    float* f = D->tmp0.as<float>(n);
    float* o = D->tmp1.as<float>(n);
    upload(D, field, f);
    ck(vflip_advect_rk2(D->ni, D->nj, D->nk, s._hx, dt, mode, (float*)D->u.p, (float*)D->v.p, (float*)D->w.p, f,
                        (int)field._nx, (int)field._ny, (int)field._nz, (float)field._ox, (float)field._oy,
                        (float)field._oz, o, nullptr, nullptr, nullptr),
       "advect_rk2");
    download(D, o, out);
}
}  // namespace
// Grid advection: RK2 backtrace and trilinear gather (SubDomainSolver::advect_field,
// S/sub_domain.cpp:191-235; the coarse grid's own RK3 form, S/Smoke_solver3D.cpp:564-656,
// is not part of its live step)
void SmokeSolver3D::advect_field(float dt, buffer3Df& f, buffer3Df& o) { advect_mode(dev, *this, dt, 0, f, o); }
// S/Smoke_solver3D.cpp:657-707
void SmokeSolver3D::advect_field_cubic_clamp(float dt, buffer3Df& f, buffer3Df& o) {
    advect_mode(dev, *this, dt, 1, f, o);
}

void SmokeSolver3D::getDelta(buffer3Df& o, buffer3Df& n, buffer3Df& t) { delta(dev, o, n, t); }
// S/Smoke_solver3D.cpp:1320-1325
void SmokeSolver3D::compute_delta_u() {
    getDelta(_utemp, _un, _utemp);
    getDelta(_vtemp, _vn, _vtemp);
    getDelta(_wtemp, _wn, _wtemp);
}
// S/Smoke_solver3D.cpp:1327-1346
void SmokeSolver3D::hash_FLIP() {
    upload_particles(dev, FLIP, true);
    bin(dev, (int)FLIP.size(), 0, Vec3f(0.f));
}
// S/Smoke_solver3D.cpp:1451-1510
void SmokeSolver3D::particle_to_grid(buffer3Df& field, buffer3Df& coef, int component) {
    if (dev->np_binned != (int)FLIP.size()) hash_FLIP();
    upload_particles(dev, FLIP, true);
    int fx, fy, fz;
    face_dims(dev, component, fx, fy, fz);
    const size_t nf = (size_t)fx * fy * fz;
    float* f = dev->tmp0.as<float>(nf);
    float* c = dev->tmp1.as<float>(nf);
    float* F[3] = {nullptr, nullptr, nullptr};
    float* C[3] = {nullptr, nullptr, nullptr};
    F[component] = f;
    C[component] = c;
    const int np = (int)FLIP.size();
    ck(vflip_p2g(dev->P, VFLIP_VARIANT_COARSE, exact_transfers ? 1 : 0, np, np ? (float*)dev->pos.p : nullptr,
                 np ? (float*)dev->vel.p : nullptr, (int*)dev->start.p, (int*)dev->order.p, 0.f, 0.f, 0.f, F[0], C[0],
                 F[1], C[1], F[2], C[2]),
       "p2g");
    download(dev, f, field);
    download(dev, c, coef);
}
// S/Smoke_solver3D.cpp:1398-1406: vel += trilinear(_utemp, _vtemp, _wtemp)
void SmokeSolver3D::particle_from_grid() {
    const int np = (int)FLIP.size();
    if (!np) return;
    upload_particles(dev, FLIP, true);
    float* du = dev->tmp0.as<float>(dev->nu());
    float* dv = dev->tmp1.as<float>(dev->nv());
    float* dw = dev->tmp2.as<float>(dev->nw());
    upload(dev, _utemp, du);
    upload(dev, _vtemp, dv);
    upload(dev, _wtemp, dw);
    ck(vflip_g2p(dev->ni, dev->nj, dev->nk, _hx, np, (float*)dev->pos.p, (float*)dev->vel.p, du, dv, dw, nullptr,
                 nullptr, nullptr, 1.0f, 0.f, 0.f, 0.f, nullptr),
       "g2p");
    download_particles(nullptr, (float*)dev->vel.p, (size_t)np, FLIP);
}
// S/Smoke_solver3D.cpp:1407-1430 (traceRK3 through get_velocity, solid push-out;
// with a subdomain the blended velocity and its solid, S/Smoke_solver3D.cpp:71-124)
void SmokeSolver3D::advect_FLIP(float dt) {
    const int np = (int)FLIP.size();
    if (!np) return;
    upload_particles(dev, FLIP, false);
    upload_flow(dev, *this);
    if (sub_domain.dev) {
        upload_flow(sub_domain.dev, sub_domain);
        const vflip_coarse_view C = coarse_view(dev);
        const vflip_subdomain_view S = sub_view(sub_domain.dev, sub_domain);
        ck(vflip_advect_particles_nested(&C, &S, dt, np, (float*)dev->pos.p, nullptr), "advect_particles_nested");
    } else {
        ck(vflip_advect_particles(dev->ni, dev->nj, dev->nk, _hx, dt, VFLIP_VARIANT_COARSE, 0.f, 0.f, 0.f, np,
                                  (float*)dev->pos.p, (float*)dev->u.p, (float*)dev->v.p, (float*)dev->w.p,
                                  (float*)dev->phi.p, nullptr),
           "advect_particles");
    }
    download_particles((float*)dev->pos.p, nullptr, (size_t)np, FLIP);
    dev->np_binned = -1;
}
// S/Smoke_solver3D.cpp:1347-1396 (rand() jitter replaced by the counter hash)
void SmokeSolver3D::reorder_FLIP() {
    hash_FLIP();
    upload_flow(dev, *this);
    const int np = (int)FLIP.size();
    const int n = select_into(dev, np + 8 * (int)dev->nc(), [&](int cap, float* op, float* ov, int* nout) {
        return vflip_reorder(dev->P, VFLIP_VARIANT_COARSE, 0.f, 0.f, 0.f, seed, np, np ? (float*)dev->pos.p : nullptr,
                             np ? (float*)dev->vel.p : nullptr, (int*)dev->start.p,
                             np ? (int*)dev->order.p : nullptr, (float*)dev->u.p, (float*)dev->v.p, (float*)dev->w.p,
                             (float*)dev->phi.p, cap, op, ov, nout);
    });
    ++seed;
    download_particles((float*)dev->pos2.p, (float*)dev->vel2.p, (size_t)n, FLIP);
    dev->np_binned = -1;
}
// S/Smoke_solver3D.cpp:1529-1560: ghosts appended to the subdomain's particles
void SmokeSolver3D::seed_ghost() {
    if (!sub_domain.dev) return;
    DeviceState* D = sub_domain.dev;
    upload_flow(dev, *this);
    upload_flow(D, sub_domain);
    const vflip_coarse_view C = coarse_view(dev);
    const vflip_subdomain_view S = sub_view(D, sub_domain);
    const int n = select_into(D, 8 * (int)D->nc(), [&](int cap, float* op, float* ov, int* nout) {
        return vflip_seed_ghost(D->P, &C, &S, seed, cap, op, ov, nout);
    });
    ++seed;
    std::vector<particle> g;
    download_particles((float*)D->pos2.p, (float*)D->vel2.p, (size_t)n, g);
    sub_domain.FLIP.insert(sub_domain.FLIP.end(), g.begin(), g.end());
}
// S/Smoke_solver3D.cpp:1560-1621: PSE viscosity on the subdomain's particles (ghosts
// seeded, binned with them, then dropped again)
void SmokeSolver3D::diffuse(float dt) {
    if (!sub_domain.dev) return;
    const int num = (int)sub_domain.FLIP.size();
    seed_ghost();
    DeviceState* D = sub_domain.dev;
    sub_domain.hash_FLIP();
    const vflip_coarse_view C = coarse_view(dev);
    const vflip_subdomain_view S = sub_view(D, sub_domain);
    float* out = D->tmp0.as<float>(3 * (size_t)std::max(num, 1));
    if (num)
        ck(vflip_pse_diffuse(&C, &S, dt, nu, num, (float*)D->pos.p, (float*)D->vel.p, (int*)D->start.p,
                             (int*)D->order.p, out, nullptr),
           "pse_diffuse");
    sub_domain.FLIP.resize(num);
    download_particles(nullptr, out, (size_t)num, sub_domain.FLIP);
    D->np_binned = -1;
}
// S/Smoke_solver3D.h:47-96 (rand() replaced by the counter hash)
void SmokeSolver3D::emit_tracers() {
    std::vector<Vec4f> kept;
    for (const Vec4f& t : tracers)
        if (t[0] > 2 * _hx && t[1] > 2 * _hx && t[2] > 2 * _hx && t[0] < _lx - 2 * _hx && t[1] < _ly - 2 * _hx &&
            t[2] < _lz - 2 * _hx && t[3] > 0.01)
            kept.push_back(t);
    tracers.swap(kept);
    auto append = [&](const float* d, int n) {
        std::vector<float> h(4 * (size_t)std::max(n, 1));
        if (n) hk(hipMemcpy(h.data(), d, 16 * (size_t)n, hipMemcpyDeviceToHost), "download tracers");
        for (int q = 0; q < n; ++q) tracers.push_back(Vec4f(h[4 * q], h[4 * q + 1], h[4 * q + 2], h[4 * q + 3]));
    };
    if (emitter_n) {
        float* d = dev->tr.as<float>(4 * (size_t)emitter_n);
        ck(vflip_emit_tracers_sphere(dev->P, seed, (int)emitter_n, emitter_pos[0], emitter_pos[1], emitter_pos[2],
                                     emitter_r, _hx, d),
           "emit_tracers_sphere");
        append(d, (int)emitter_n);
    }
    if (sub_domain.dev) {
        DeviceState* D = sub_domain.dev;
        upload_flow(dev, *this);
        upload_flow(D, sub_domain);
        const vflip_coarse_view C = coarse_view(dev);
        const vflip_subdomain_view S = sub_view(D, sub_domain);
        const int cap = 4 * (int)D->nc();
        float* d = D->tr.as<float>(4 * (size_t)cap);
        int n = 0;
        ck(vflip_emit_tracers_sub(D->P, &C, &S, seed, cap, d, &n), "emit_tracers_sub");
        append(d, n);
    }
    ++seed;
}
// S/Smoke_solver3D.h:101-125
void SmokeSolver3D::advect_tracers(float dt) {
    const int n = (int)tracers.size();
    if (!n) return;
    float* d = dev->tr.as<float>(4 * (size_t)n);
    hk(hipMemcpy(d, tracers.data(), 16 * (size_t)n, hipMemcpyHostToDevice), "upload tracers");
    upload_flow(dev, *this);
    const vflip_coarse_view C = coarse_view(dev);
    if (sub_domain.dev) {
        upload_flow(sub_domain.dev, sub_domain);
        const vflip_subdomain_view S = sub_view(sub_domain.dev, sub_domain);
        ck(vflip_advect_tracers(&C, &S, dt, _dens_decay, n, d, nullptr), "advect_tracers");
    } else {
        ck(vflip_advect_tracers(&C, nullptr, dt, _dens_decay, n, d, nullptr), "advect_tracers");
    }
    hk(hipMemcpy(tracers.data(), d, 16 * (size_t)n, hipMemcpyDeviceToHost), "download tracers");
}
// S/Smoke_solver3D.h:126-161
void SmokeSolver3D::write_tracers(const char* file_path, int frame) {
    const int n = (int)tracers.size();
    std::vector<float> rec(4 * (size_t)std::max(n, 1));
    if (n) {
        float* d = dev->tr.as<float>(8 * (size_t)n);
        hk(hipMemcpy(d, tracers.data(), 16 * (size_t)n, hipMemcpyHostToDevice), "upload tracers");
        ck(vflip_tracer_frame(n, d, _lx, d + 4 * (size_t)n, nullptr), "tracer_frame");
        hk(hipMemcpy(rec.data(), d + 4 * (size_t)n, 16 * (size_t)n, hipMemcpyDeviceToHost), "download frame");
    }
    char name[1024];
    std::snprintf(name, sizeof name, "%s/Particle_data%04d.bin", file_path, frame);
    FILE* f = std::fopen(name, "wb");
    if (!f) fail("write_tracers: cannot open the frame file", -1);
    std::fwrite(rec.data(), sizeof(float), 4 * (size_t)n, f);
    std::fclose(f);
}
// S/Smoke_solver3D.cpp:32-58: the largest face speed (signed, as the reference keeps it)
void SmokeSolver3D::getCFL(float /*dt*/) {
    float max_v = _hx;
    for (uint k = 0; k < _nz; ++k)
        for (uint j = 0; j < _ny; ++j)
            for (uint i = 0; i < _nx + 1; ++i)
                if (std::fabs(_un(i, j, k)) > max_v) max_v = _un(i, j, k);
    for (uint k = 0; k < _nz; ++k)
        for (uint j = 0; j < _ny + 1; ++j)
            for (uint i = 0; i < _nx; ++i)
                if (std::fabs(_vn(i, j, k)) > max_v) max_v = _vn(i, j, k);
    for (uint k = 0; k < _nz + 1; ++k)
        for (uint j = 0; j < _ny; ++j)
            for (uint i = 0; i < _nx; ++i)
                if (std::fabs(_wn(i, j, k)) > max_v) max_v = _wn(i, j, k);
    _cfl = _hx / max_v;
}
// S/Smoke_solver3D.cpp:313-411: the nested FLIP step
void SmokeSolver3D::time_step(float dt, int /*adv_type*/, int /*frame*/) {
    getCFL(dt);
    float t = 0;
    emit_tracers();
    while (t < dt) {
        float substep = 100 * _cfl;
        if (t + substep > dt) substep = dt - t;
        if (!(substep > 0)) break;  // the reference's signed CFL can give a negative substep
        advect_tracers(substep);
        _utemp.setZero();
        _vtemp.setZero();
        _wtemp.setZero();
        diffuse(substep);
        FLIP.insert(FLIP.end(), sub_domain.FLIP.begin(), sub_domain.FLIP.end());
        advect_FLIP(substep);
        hash_FLIP();
        particle_to_grid(_un, _unp1, 0);
        particle_to_grid(_vn, _vnp1, 1);
        particle_to_grid(_wn, _wnp1, 2);
        _utemp.copy(_un);
        _vtemp.copy(_vn);
        _wtemp.copy(_wn);
        pcg_projection(1.0f);
        compute_delta_u();
        particle_from_grid();
        if (sub_domain.dev) {
            sub_domain.FLIP.clear();
            for (particle& p : FLIP)
                if (sub_domain.is_halo(p.pos, 0.5f * _hx)) sub_domain.FLIP.push_back(p);
            sub_domain.hash_FLIP();
            sub_domain.particle_to_grid();
            sub_domain.pcg_projection(1.0f);
            sub_domain.compute_delta_u();
            sub_domain.particle_from_grid();
            sub_domain.reorder_FLIP();
        }
        reorder_FLIP();
        t += substep;
    }
}

// ================================================= CoarseFineSmokeSolver
CoarseFineSmokeSolver::CoarseFineSmokeSolver() { _N = 8; }
CoarseFineSmokeSolver::~CoarseFineSmokeSolver() { Finalize(); }
void CoarseFineSmokeSolver::Finalize() {
    fine_solver.Finalize();
    coarse_solver.Finalize();
}
// S/coarse_fine_solver.cpp:120-128 initialises the coarse solver twice and never the
// fine one; the evident intent is kept: fine grid nx x ny x nz, coarse grid 1/_N of it.
void CoarseFineSmokeSolver::initialize(uint nx, uint ny, uint nz, float L) {
    _nx = nx; _ny = ny; _nz = nz;
    _h = L / (float)_nx;
    fine_solver.init(nx, ny, nz, L, 1.0f);
    coarse_solver.init(nx / _N, ny / _N, nz / _N, L, 1.0f);
}
void CoarseFineSmokeSolver::FineProjection() { fine_solver.pcg_projection(1.0f); }

namespace {
struct Pair {
    DBuf f, c, d;
};
}  // namespace
// S/coarse_fine_solver.cpp:182-230
void CoarseFineSmokeSolver::SetFaceBoundary(buffer3Df& fine, buffer3Df& coarse, int offx, int offy, int offz) {
    DeviceState* D = fine_solver.dev;
    Pair b;
    float* f = b.f.as<float>((size_t)fine._nx * fine._ny * fine._nz);
    float* c = b.c.as<float>((size_t)coarse._nx * coarse._ny * coarse._nz);
    upload(D, fine, f);
    upload(D, coarse, c);
    ck(vflip_set_face_boundary(f, (int)fine._nx, (int)fine._ny, (int)fine._nz, c, (int)coarse._nx, (int)coarse._ny,
                               (int)coarse._nz, (int)_N, offx, offy, offz, nullptr),
       "set_face_boundary");
    download(D, f, fine);
}
// S/coarse_fine_solver.cpp:73-108
void CoarseFineSmokeSolver::FineToCoarseOperator(buffer3Df& fine, buffer3Df& coarse, int offx, int offy, int offz) {
    DeviceState* D = fine_solver.dev;
    Pair b;
    float* f = b.f.as<float>((size_t)fine._nx * fine._ny * fine._nz);
    float* c = b.c.as<float>((size_t)coarse._nx * coarse._ny * coarse._nz);
    upload(D, fine, f);
    ck(vflip_fine_to_coarse(f, (int)fine._nx, (int)fine._ny, (int)fine._nz, c, (int)coarse._nx, (int)coarse._ny,
                            (int)coarse._nz, (int)_N, offx, offy, offz, nullptr),
       "fine_to_coarse");
    download(D, c, coarse);
}
// S/coarse_fine_solver.cpp:110-118 (velocity faces with the offset _N/2 along their normal;
// the density is on the reference's doubled grid and is averaged with offset 0)
void CoarseFineSmokeSolver::FineToCoarse() {
    FineToCoarseOperator(fine_solver._un, coarse_solver._un, (int)_N / 2, 0, 0);
    FineToCoarseOperator(fine_solver._vn, coarse_solver._vn, 0, (int)_N / 2, 0);
    FineToCoarseOperator(fine_solver._wn, coarse_solver._wn, 0, 0, (int)_N / 2);
    FineToCoarseOperator(fine_solver._rho, coarse_solver._rho, 0, 0, 0);
    FineToCoarseOperator(fine_solver._Temperature, coarse_solver._Temperature, 0, 0, 0);
}
namespace {
void coarse_to_fine(DeviceState* D, buffer3Df& fine, buffer3Df& coarse, buffer3Df* minus, buffer3Df& out) {
    Pair b;
    const size_t nf = (size_t)fine._nx * fine._ny * fine._nz;
    float* f = b.f.as<float>(nf);
    float* c = b.c.as<float>((size_t)coarse._nx * coarse._ny * coarse._nz);
    float* m = nullptr;
    upload(D, coarse, c);
    if (minus) {
        m = b.d.as<float>(nf);
        upload(D, *minus, m);
    }
    ck(vflip_coarse_to_fine(f, (int)fine._nx, (int)fine._ny, (int)fine._nz, (float)fine._ox, (float)fine._oy,
                            (float)fine._oz, (float)fine._hx, c, (int)coarse._nx, (int)coarse._ny, (int)coarse._nz,
                            (float)coarse._ox, (float)coarse._oy, (float)coarse._oz, (float)coarse._hx, m, nullptr),
       "coarse_to_fine");
    download(D, f, out);
}
}  // namespace
// S/coarse_fine_solver.cpp:146-170
void CoarseFineSmokeSolver::CoarseToFineOperator(buffer3Df& fine, buffer3Df& coarse) {
    coarse_to_fine(fine_solver.dev, fine, coarse, nullptr, fine);
}
// S/coarse_fine_solver.cpp:47-71
void CoarseFineSmokeSolver::ComputeSmallDifference(buffer3Df& fine, buffer3Df& coarse, buffer3Df& diff) {
    coarse_to_fine(fine_solver.dev, fine, coarse, &fine, diff);
}
// S/coarse_fine_solver.cpp:35-44
void CoarseFineSmokeSolver::LargeScaleAdvection(float dt) {
    coarse_solver.advect_field(dt, coarse_solver._un, coarse_solver._unp1);
    coarse_solver.advect_field(dt, coarse_solver._vn, coarse_solver._vnp1);
    coarse_solver.advect_field(dt, coarse_solver._wn, coarse_solver._wnp1);
    CoarseToFineOperator(fine_solver._unp1, coarse_solver._unp1);
    CoarseToFineOperator(fine_solver._vnp1, coarse_solver._vnp1);
    CoarseToFineOperator(fine_solver._wnp1, coarse_solver._wnp1);
}

// ------------------------------------------------------------ MultiGridSolver3D
// Device side of one MultiGridSolver3D: a vflip_mg3d handle per level it has been
// asked to V-cycle from (the levels of a level's grid are the finer grid's levels
// below it) and scratch copies of the fields.
struct MGDevice {
    int type = VFLIP_MG3D_FLOAT;
    int dirichlet = -1;
    std::vector<vflip_mg3d*> handles;
    DBuf x, b, t;
    ~MGDevice() {
        for (auto* h : handles)
            if (h) vflip_mg3d_destroy(h);
    }
    vflip_mg3d* handle(int level, const uniform_grid_descriptor_3D& g, int dir) {
        if (dir != dirichlet) {
            for (auto*& h : handles)
                if (h) { vflip_mg3d_destroy(h); h = nullptr; }
            dirichlet = dir;
        }
        if ((int)handles.size() <= level) handles.resize(level + 1, nullptr);
        if (!handles[level]) ck(vflip_mg3d_create(g.gridx, g.gridy, g.gridz, type, dir, &handles[level]), "mg3d_create");
        return handles[level];
    }
};

namespace {
template <class T> int mg_type();
template <> int mg_type<float>() { return VFLIP_MG3D_FLOAT; }
template <> int mg_type<double>() { return VFLIP_MG3D_DOUBLE; }

// Buffer3D bricks <-> a linear x-fastest device array of g's extents.
template <class T> T* mg_upload(DBuf& d, const Buffer3D<T>& b, const uniform_grid_descriptor_3D& g) {
    const size_t n = (size_t)g.gridx * g.gridy * g.gridz;
    std::vector<T> lin(n);
    for (int k = 0; k < g.gridz; ++k)
        for (int j = 0; j < g.gridy; ++j)
            for (int i = 0; i < g.gridx; ++i) lin[i + (size_t)g.gridx * (j + (size_t)g.gridy * k)] = b(i, j, k);
    T* p = d.as<T>(n);
    hk(hipMemcpy(p, lin.data(), n * sizeof(T), hipMemcpyHostToDevice), "mg upload");
    return p;
}
template <class T> void mg_download(const T* p, Buffer3D<T>& b, const uniform_grid_descriptor_3D& g) {
    const size_t n = (size_t)g.gridx * g.gridy * g.gridz;
    std::vector<T> lin(n);
    hk(hipMemcpy(lin.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "mg download");
    for (int k = 0; k < g.gridz; ++k)
        for (int j = 0; j < g.gridy; ++j)
            for (int i = 0; i < g.gridx; ++i) b(i, j, k) = lin[i + (size_t)g.gridx * (j + (size_t)g.gridy * k)];
}
uniform_grid_descriptor_3D mg_coarser(const uniform_grid_descriptor_3D& g) {
    return uniform_grid_descriptor_3D(std::max(g.gridx / 2, 1), std::max(g.gridy / 2, 1), std::max(g.gridz / 2, 1));
}
void mg_same(const uniform_grid_descriptor_3D& a, const uniform_grid_descriptor_3D& b, const char* what) {
    if (a.gridx != b.gridx || a.gridy != b.gridy || a.gridz != b.gridz)
        throw std::runtime_error(std::string("vflip facade: ") + what + ": the coarser level must be max(n/2, 1)");
}
}  // namespace

// S/Multigrid3D.h:124-137
template <class T> void MultiGrid<T>::m_ComputeLevels() {
    int level = 0;
    int x = m_system_descriptor.gridx, y = m_system_descriptor.gridy, z = m_system_descriptor.gridz;
    while ((size_t)x * y * z > 512) {
        ++level;
        x = std::max(x / 2, 1);
        y = std::max(y / 2, 1);
        z = std::max(z / 2, 1);
    }
    m_max_level = level + 1;
}
// S/Multigrid3D.h:71-118: x, b, r and x_new for the coarse levels, x_new and r for the finest
template <class T> void MultiGrid<T>::m_AssignMemoryEachLevel() {
    xk.assign(m_max_level, nullptr); xk_new.assign(m_max_level, nullptr);
    bk.assign(m_max_level, nullptr); rk.assign(m_max_level, nullptr);
    systemk.assign(m_max_level, uniform_grid_descriptor_3D());
    systemk[0] = m_system_descriptor;
    for (int i = 1; i < m_max_level; ++i) systemk[i] = mg_coarser(systemk[i - 1]);
    auto make = [&](int l, bool init) {
        auto* b = new Buffer3D<T>();
        if (init) b->init(systemk[l].gridx, systemk[l].gridy, systemk[l].gridz, 1.0 / systemk[l].gridx, -0.5, -0.5, -0.5);
        owned_.push_back(b);
        return b;
    };
    for (int l = 0; l < m_max_level; ++l) {
        xk[l] = make(l, l > 0);
        bk[l] = make(l, l > 0);
        xk_new[l] = make(l, true);
        rk[l] = make(l, true);
    }
    m_bInitialized = true;
}
template <class T> void MultiGrid<T>::m_FinalMemoryEachLevel() {
    for (auto* b : owned_) delete b;
    owned_.clear();
    xk.clear(); xk_new.clear(); bk.clear(); rk.clear();
    m_bInitialized = false;
}
// S/Multigrid3D.h:138-150
template <class T> void MultiGrid<T>::m_InitialSystem(int gridx, int gridy, int gridz) {
    m_FinalMemoryEachLevel();
    m_SetGridInformation(gridx, gridy, gridz);
    m_ComputeLevels();
    m_AssignMemoryEachLevel();
}

template <class T> MultiGridSolver3D<T>::~MultiGridSolver3D() { delete dev_; }
template <class T> MGDevice& MultiGridSolver3D<T>::dev() {
    if (!dev_) {
        dev_ = new MGDevice();
        dev_->type = mg_type<T>();
    }
    return *dev_;
}
// S/Multigrid3D.h:1328-1392; the coarse levels' host buffers are left zero, as there
template <class T> void MultiGridSolver3D<T>::m_Vcycle(Buffer3D<T>* x, Buffer3D<T>* b, T, T&, int level) {
    if (level < 0 || level >= this->m_max_level) throw std::runtime_error("vflip facade: m_Vcycle level");
    this->xk[level] = x;
    this->bk[level] = b;
    MGDevice& D = dev();
    const auto& g = this->systemk[level];
    vflip_mg3d* h = D.handle(level, g, _b_Dirichlet ? 1 : 0);
    T* dx = mg_upload(D.x, *x, g);
    T* db = mg_upload(D.b, *b, g);
    ck(vflip_mg3d_vcycle(h, dx, db, nullptr), "mg3d_vcycle");
    mg_download(dx, *x, g);
}
// S/Multigrid3D.h:1394-1436
template <class T> void MultiGridSolver3D<T>::m_FullMultiGrid(Buffer3D<T>* x, Buffer3D<T>* b, T, T&) {
    this->xk[0] = x;
    this->bk[0] = b;
    MGDevice& D = dev();
    const auto& g = this->systemk[0];
    vflip_mg3d* h = D.handle(0, g, _b_Dirichlet ? 1 : 0);
    T* dx = mg_upload(D.x, *x, g);
    T* db = mg_upload(D.b, *b, g);
    ck(vflip_mg3d_full_multigrid(h, dx, db, nullptr), "mg3d_full_multigrid");
    mg_download(dx, *x, g);
}
// S/Multigrid3D.h:840-1129 (x_new is untouched here: the device keeps its own scratch)
template <class T>
void MultiGridSolver3D<T>::m_Red_Black_Gauss(uniform_grid_descriptor_3D& g, Buffer3D<T>& b, Buffer3D<T>& x,
                                             Buffer3D<T>&, int iter_time, double omega) {
    MGDevice& D = dev();
    T* dx = mg_upload(D.x, x, g);
    T* db = mg_upload(D.b, b, g);
    T* dt = D.t.as<T>((size_t)g.gridx * g.gridy * g.gridz);
    ck(vflip_mg3d_rbgs(g.gridx, g.gridy, g.gridz, mg_type<T>(), _b_Dirichlet ? 1 : 0, db, dx, dt, iter_time, omega,
                       nullptr), "mg3d_rbgs");
    mg_download(dx, x, g);
}
// S/Multigrid3D.h:1130-1186
template <class T>
void MultiGridSolver3D<T>::m_ComputeResidual(uniform_grid_descriptor_3D& g, Buffer3D<T>& res, Buffer3D<T>& b,
                                             Buffer3D<T>& x) {
    MGDevice& D = dev();
    T* dx = mg_upload(D.x, x, g);
    T* db = mg_upload(D.b, b, g);
    T* dr = D.t.as<T>((size_t)g.gridx * g.gridy * g.gridz);
    ck(vflip_mg3d_residual(g.gridx, g.gridy, g.gridz, mg_type<T>(), _b_Dirichlet ? 1 : 0, dr, db, dx, nullptr),
       "mg3d_residual");
    mg_download(dr, res, g);
}
// S/Multigrid3D.h:1187-1191
template <class T> void MultiGridSolver3D<T>::m_ExactSolve() {
    const int l = this->m_max_level - 1;
    m_Red_Black_Gauss(this->systemk[l], *this->bk[l], *this->xk[l], *this->xk_new[l], 100, 1.0);
}
// S/Multigrid3D.h:1193-1245
template <class T>
void MultiGridSolver3D<T>::m_Restrict(uniform_grid_descriptor_3D& next_level, uniform_grid_descriptor_3D& curr_level,
                                      Buffer3D<T>& src, Buffer3D<T>& dst) {
    mg_same(next_level, mg_coarser(curr_level), "m_Restrict");
    MGDevice& D = dev();
    T* ds = mg_upload(D.x, src, curr_level);
    T* dd = D.t.as<T>((size_t)next_level.gridx * next_level.gridy * next_level.gridz);
    ck(vflip_mg3d_restrict(curr_level.gridx, curr_level.gridy, curr_level.gridz, mg_type<T>(), ds, dd, nullptr),
       "mg3d_restrict");
    mg_download(dd, dst, next_level);
}
// S/Multigrid3D.h:1298-1326 (next_level: the finer grid of dst, curr_level: src's)
template <class T>
void MultiGridSolver3D<T>::m_Prolongate(uniform_grid_descriptor_3D& next_level, uniform_grid_descriptor_3D& curr_level,
                                        Buffer3D<T>& src, Buffer3D<T>& dst) {
    mg_same(curr_level, mg_coarser(next_level), "m_Prolongate");
    MGDevice& D = dev();
    T* ds = mg_upload(D.b, src, curr_level);
    T* dd = mg_upload(D.x, dst, next_level);
    ck(vflip_mg3d_prolongate(next_level.gridx, next_level.gridy, next_level.gridz, mg_type<T>(), ds, dd, nullptr),
       "mg3d_prolongate");
    mg_download(dd, dst, next_level);
}

template class MultiGrid<float>;
template class MultiGrid<double>;
template class MultiGridSolver3D<float>;
template class MultiGridSolver3D<double>;

}  // namespace ivock
