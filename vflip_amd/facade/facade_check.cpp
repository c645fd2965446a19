// Drives the ivock facade the way reference code drives SmokeSolver3D and
// AMGPCGSolve, and writes inputs and outputs as raw little-endian arrays for
// tests/test_gpu_facade.py to check against the oracle:
//   1. SmokeSolver3D::init, set_boundary_phi (box walls + sphere), set_boundary_vel,
//      a seeded MAC velocity, pcg_projection;
//   2. the FLIP particles init() seeded: velocities set, hash_FLIP, particle_to_grid
//      for the three components (the reference's summation order);
//   3. ivock::AMGPCGSolve<double> on a SparseMatrixd holding a 7-point cut-cell system;
//   4. MGSolverf / MGSolverd (MultiGridSolver3D) V-cycles, full multigrid and level operators.
// Usage: facade_check <output directory>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "vflip_ivock.hpp"

using namespace ivock;

namespace {
std::string dir;
template <class T> void save(const char* name, const std::vector<T>& a) {
    FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
    if (!f) { std::perror(name); std::exit(2); }
    std::fwrite(a.data(), sizeof(T), a.size(), f);
    std::fclose(f);
}
// linear x-fastest copy of a brick buffer (the brick index goes through operator())
std::vector<float> linear(const buffer3Df& b) {
    std::vector<float> out((size_t)b._nx * b._ny * b._nz);
    for (uint k = 0; k < b._nz; ++k)
        for (uint j = 0; j < b._ny; ++j)
            for (uint i = 0; i < b._nx; ++i) out[i + (size_t)b._nx * (j + (size_t)b._ny * k)] = b(i, j, k);
    return out;
}
uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
float urand(uint32_t key) { return (float)(hash32(key) >> 8) / 16777216.0f * 2.0f - 1.0f; }

const int N = 24;
float sdf(const Vec3f& p) {
    const float h = 1.0f / N;
    float d = std::fmin(std::fmin(p[0] - h, (N - 2) * h - p[0]), std::fmin(p[1] - h, std::fmin(p[2] - h, (N - 2) * h - p[2])));
    const float s = std::sqrt((p[0] - 0.45f) * (p[0] - 0.45f) + (p[1] - 0.5f) * (p[1] - 0.5f) + (p[2] - 0.55f) * (p[2] - 0.55f)) - 0.17f;
    return std::fmin(d, s);
}
Vec3f still(const Vec3f&) { return Vec3f(0.f, 0.f, 0.f); }
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: facade_check <dir>\n"); return 2; }
    dir = argv[1];
    SmokeSolver3D s;
    s.init(N, N, N, 1.0f, 1.0f);
    std::printf("init: %zu FLIP particles\n", s.FLIP.size());
    s.set_boundary_phi(sdf);
    s.set_boundary_vel(still);
    buffer3Df* vel[3] = {&s._un, &s._vn, &s._wn};
    for (int c = 0; c < 3; ++c)
        for (uint k = 0; k < vel[c]->_nz; ++k)
            for (uint j = 0; j < vel[c]->_ny; ++j)
                for (uint i = 0; i < vel[c]->_nx; ++i)
                    (*vel[c])(i, j, k) = 0.5f * urand((uint32_t)(c * 1000003 + i + 131 * j + 17161 * k));
    save("phi.f32", linear(s._nodal_solid_phi));
    save("u0.f32", linear(s._un));
    save("v0.f32", linear(s._vn));
    save("w0.f32", linear(s._wn));
    s.pcg_projection(1.0f);
    std::printf("projection: %d iterations, rel residual %.3e, FP64 true residual %.3e\n", s.last_stats.iterations,
                s.last_stats.rel_residual, s.last_stats.true_residual);
    save("uw.f32", linear(s._u_weights));
    save("vw.f32", linear(s._v_weights));
    save("ww.f32", linear(s._w_weights));
    save("u1.f32", linear(s._un));
    save("v1.f32", linear(s._vn));
    save("w1.f32", linear(s._wn));
    save("p.f64", s.pressure);
    save("rhs.f64", s.rhs);
    save("stats.f64", std::vector<double>{(double)s.last_stats.iterations, s.last_stats.rel_residual,
                                          s.last_stats.true_residual});

    // FLIP transfer: the init()-seeded particles with seeded velocities
    for (size_t q = 0; q < s.FLIP.size(); ++q)
        for (int a = 0; a < 3; ++a) s.FLIP[q].vel[a] = urand((uint32_t)(7000001 + 3 * q + a));
    std::vector<float> pos(3 * s.FLIP.size()), pv(3 * s.FLIP.size());
    for (size_t q = 0; q < s.FLIP.size(); ++q)
        for (int a = 0; a < 3; ++a) { pos[3 * q + a] = s.FLIP[q].pos[a]; pv[3 * q + a] = s.FLIP[q].vel[a]; }
    save("fpos.f32", pos);
    save("fvel.f32", pv);
    s.hash_FLIP();
    s.particle_to_grid(s._un, s._unp1, 0);
    s.particle_to_grid(s._vn, s._vnp1, 1);
    s.particle_to_grid(s._wn, s._wnp1, 2);
    save("pu.f32", linear(s._un));
    save("pv.f32", linear(s._vn));
    save("pw.f32", linear(s._wn));
    save("cu.f32", linear(s._unp1));
    save("cv.f32", linear(s._vnp1));
    save("cw.f32", linear(s._wnp1));

    // AMGPCGSolve<double> on the 7-point system of S/Smoke_solver3D.cpp:1012-1096 shape
    const int ni = 20, nj = 18, nk = 16;
    const unsigned n = (unsigned)(ni * nj * nk);
    SparseMatrixd A(n);
    std::vector<double> b(n, 0.0), x;
    std::vector<double> d(n, 0), ax(n, 0), ay(n, 0), az(n, 0);
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i) {
                const unsigned c = (unsigned)(i + ni * (j + nj * k));
                double diag = 0;
                const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
                for (auto& o : nb) {
                    const int a = i + o[0], bb = j + o[1], e = k + o[2];
                    // face weight in (0.25, 1], a single FP32 term as the reference's
                    // assembly writes it (S/Smoke_solver3D.cpp:1012-1096): a seeded cut-cell system
                    const unsigned lo = std::min(c, (unsigned)(a + ni * (bb + nj * e)));
                    const double w = (double)(0.625f + 0.375f * urand(9000011u + 6 * lo + (o[0] ? 0 : o[1] ? 1 : 2)));
                    diag += w;
                    if (a >= 1 && a < ni - 1 && bb >= 1 && bb < nj - 1 && e >= 1 && e < nk - 1)
                        A.set_element(c, (unsigned)(a + ni * (bb + nj * e)), -w);
                }
                A.set_element(c, c, diag);
                b[c] = urand(12000017u + c);
            }
    for (unsigned r = 0; r < n; ++r)
        for (size_t q = 0; q < A.index[r].size(); ++q) {
            const unsigned c = A.index[r][q];
            if (c == r) d[r] = A.value[r][q];
            else if (c == r + 1) ax[r] = A.value[r][q];
            else if (c == r + (unsigned)ni) ay[r] = A.value[r][q];
            else if (c == r + (unsigned)(ni * nj)) az[r] = A.value[r][q];
        }
    double res = 0;
    int its = 0;
    const bool ok = AMGPCGSolve<double>(A, b, x, 1e-6, 50, res, its, ni, nj, nk);
    std::printf("AMGPCGSolve: ok=%d iterations=%d residual=%.3e\n", (int)ok, its, res);
    save("amg_d.f64", d);
    save("amg_ax.f64", ax);
    save("amg_ay.f64", ay);
    save("amg_az.f64", az);
    save("amg_b.f64", b);
    save("amg_x.f64", x);
    save("amg_stats.f64", std::vector<double>{(double)ok, (double)its, res});

    // 4. MGSolverf / MGSolverd: m_InitialSystem, two m_Vcycle calls, m_FullMultiGrid, the
    //    level operators on the finest level, Dirichlet walls.
    {
        const int gx = 20, gy = 18, gz = 24;
        MGSolverf mf;
        mf._b_Dirichlet = true;
        mf.m_InitialSystem(gx, gy, gz);
        MGSolverd md;
        md.m_InitialSystem(gx, gy, gz);
        buffer3Df bf, xf, xf2, rf, cf;
        Buffer3D<double> bd, xd;
        for (auto* q : {&bf, &xf, &xf2, &rf}) q->init(gx, gy, gz);
        cf.init(gx / 2, gy / 2, gz / 2);
        bd.init(gx, gy, gz);
        xd.init(gx, gy, gz);
        std::vector<float> lb, lx;
        for (int k = 0; k < gz; ++k)
            for (int j = 0; j < gy; ++j)
                for (int i = 0; i < gx; ++i) {
                    const uint32_t c = (uint32_t)(i + gx * (j + gy * k));
                    bf(i, j, k) = urand(15000001u + c);
                    xf(i, j, k) = 0.1f * urand(16000003u + c);
                    bd(i, j, k) = bf(i, j, k);
                    xd(i, j, k) = xf(i, j, k);
                }
        xf2.copy(xf);
        save("mg_b.f32", linear(bf));
        save("mg_x0.f32", linear(xf));
        std::printf("MGSolverf: %d levels, coarsest %dx%dx%d\n", mf.m_max_level, mf.systemk.back().gridx,
                    mf.systemk.back().gridy, mf.systemk.back().gridz);
        float tol = 0.f, resid = 0.f;
        mf.m_ComputeResidual(mf.systemk[0], rf, bf, xf);
        save("mg_res.f32", linear(rf));
        mf.m_Restrict(mf.systemk[1], mf.systemk[0], rf, cf);
        save("mg_restrict.f32", linear(cf));
        mf.m_Vcycle(&xf, &bf, tol, resid, 0);
        mf.m_Vcycle(&xf, &bf, tol, resid, 0);
        save("mg_vcycle2.f32", linear(xf));
        mf.m_FullMultiGrid(&xf2, &bf, tol, resid);
        save("mg_fmg.f32", linear(xf2));
        double told = 0, residd = 0;
        md.m_Vcycle(&xd, &bd, told, residd, 0);
        std::vector<double> outd((size_t)gx * gy * gz);
        for (int k = 0; k < gz; ++k)
            for (int j = 0; j < gy; ++j)
                for (int i = 0; i < gx; ++i) outd[i + (size_t)gx * (j + (size_t)gy * k)] = xd(i, j, k);
        save("mg_vcycle_neumann.f64", outd);
    }
    std::printf("facade check done\n");
    return 0;
}
