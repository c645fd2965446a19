// Test harness around the reference's own sources, compiled from where they lie
// in the reference checkout (never copied here). TEST INFRASTRUCTURE ONLY: it is
// built into oracle/_ref/ and used by tests/ and tests/golden/ to pin the oracle.
//
// Covered reference code (compiled as-is):
//   fraction_inside(2), fraction_inside(4)  S/levelset_util.cpp:4-103
//   Buffer3D<float>::sample_linear/_cubic/_cube_lerp/_grad   S/fluid_buffer3D.h:86-376
//   randhash                                S/util.h:238-246
//   normalize, Vec3f/Vec3i arithmetic       S/vec.h (hash_FLIP cell index,
//                                           set_boundary_* sample positions)
// The standard headers below are included first because the reference headers
// rely on the MSVC toolchain pulling them in implicitly.
#include <climits>
#include <cstring>
#include "fluid_buffer3D.h"
#include "levelset_util.h"

using ivock::Buffer3D;

extern "C" {

float ref_fraction_inside2(float a, float b) { return fraction_inside(a, b); }
float ref_fraction_inside4(float bl, float br, float tl, float tr) { return fraction_inside(bl, br, tl, tr); }
unsigned int ref_randhash(unsigned int s) { return ivock::randhash(s); }

// Loads a linear (x-fastest) field into a reference Buffer3D (8^3-blocked
// storage) and evaluates one sampler at n world positions.
// kind: 0 sample_linear, 1 sample_cubic, 2 sample_cube_lerp (corners to c8),
// 3 sample_grad (3 floats per point into out3).
void ref_sample(int kind, const float* field, int nx, int ny, int nz, double h, double ox,
                double oy, double oz, int n, const double* xyz, float* out, float* c8, float* out3) {
    Buffer3D<float> b;
    b.init(nx, ny, nz, h, ox, oy, oz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) b(i, j, k) = field[i + (size_t)nx * (j + (size_t)ny * k)];
    for (int p = 0; p < n; ++p) {
        double x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
        if (kind == 0) out[p] = b.sample_linear(x, y, z);
        else if (kind == 1) out[p] = b.sample_cubic(x, y, z);
        else if (kind == 2) {
            float* c = c8 + 8 * p;
            out[p] = b.sample_cube_lerp(x, y, z, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
        } else {
            ivock::Vec3f g = b.sample_grad(x, y, z);
            out3[3 * p] = g[0]; out3[3 * p + 1] = g[1]; out3[3 * p + 2] = g[2];
        }
    }
    b.free();
}

// normalize(Vec3f) of S/vec.h:218-219 on n xyz triples.
void ref_normalize(int n, const float* in, float* out) {
    for (int p = 0; p < n; ++p) {
        ivock::Vec3f v(in[3 * p], in[3 * p + 1], in[3 * p + 2]);
        ivock::normalize(v);
        out[3 * p] = v[0]; out[3 * p + 1] = v[1]; out[3 * p + 2] = v[2];
    }
}

// The cell-index arithmetic of hash_FLIP evaluated with the reference's own
// Vec3f/Vec3i types (S/Smoke_solver3D.cpp:1337-1339; S/sub_domain.cpp:1097-1099
// subtracts bmin first): ijk = Vec3i((pos [- bmin] + Vec3f(.5,.5,.5)*h) / h).
void ref_cell_index(int n, const float* pos, float h, int sub, float bx, float by, float bz, int* ijk) {
    const ivock::Vec3f bmin(bx, by, bz);
    for (int p = 0; p < n; ++p) {
        ivock::Vec3f xx(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        if (sub) xx = xx - bmin;
        xx += ivock::Vec3f(0.5, 0.5, 0.5) * h;
        ivock::Vec3i c(xx / h);
        ijk[3 * p] = c[0]; ijk[3 * p + 1] = c[1]; ijk[3 * p + 2] = c[2];
    }
}

// The sample positions of set_boundary_phi / set_boundary_vel evaluated with the
// reference's Vec3f arithmetic: Vec3f(i,j,k)*h - Vec3f(ox,oy,oz)*h, where the
// offsets are the Buffer3D's double _ox/_oy/_oz (S/Smoke_solver3D.cpp:806,864).
void ref_grid_points(int nx, int ny, int nz, float h, double ox, double oy, double oz, float* out) {
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                ivock::Vec3f pos = ivock::Vec3f(i, j, k) * h - ivock::Vec3f(ox, oy, oz) * h;
                float* q = out + 3 * (i + (size_t)nx * (j + (size_t)ny * k));
                q[0] = pos[0]; q[1] = pos[1]; q[2] = pos[2];
            }
}

// Blocked storage index of the reference Buffer3D for cell (i,j,k).
unsigned int ref_blocked_index(int nx, int ny, int nz, int i, int j, int k) {
    Buffer3D<float> b;
    b.init(nx, ny, nz);
    unsigned int idx = (unsigned int)(&b(i, j, k) - b.getArray()->getPtr());
    b.free();
    return idx;
}

}
