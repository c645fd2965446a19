// Coarse/fine coupling of nested subdomains and the reference's brick storage.
//
//   ComputeSmallDifference   S/coarse_fine_solver.cpp:47-71
//   FineToCoarseOperator     S/coarse_fine_solver.cpp:73-108 (FineToCoarse offsets: 110-118)
//   CoarseToFineOperator     S/coarse_fine_solver.cpp:146-170
//   SetFaceBoundary          S/coarse_fine_solver.cpp:182-230
//   Buffer3D 8^3 bricks      S/fluid_buffer3D.h:186-202
//
// Sums run in the reference's loop order (x outer, z inner) and the samplers
// are the FP64 Buffer3D ones, so results are bit-identical with the reference
// arithmetic (FMA contraction off).
#pragma clang fp contract(off)

#include "common.h"
#include "sampling.h"

namespace vflip {

// Box average of the N^3 fine cells under coarse cell (i,j,k), shifted by
// (offx,offy,offz) fine cells (N/2 along a face's normal for MAC faces).
__global__ __launch_bounds__(kBlock) void k_fine_to_coarse(const float* __restrict__ fine, int fnx, int fny, int fnz,
                                                           float* __restrict__ coarse, int cnx, int cny, int cnz,
                                                           int N, int offx, int offy, int offz) {
    const size_t c = blockIdx.x * (size_t)kBlock + threadIdx.x;
    if (c >= (size_t)cnx * cny * cnz) return;
    const int i = (int)(c % cnx), j = (int)((c / cnx) % cny), k = (int)(c / ((size_t)cnx * cny));
    int count = 0;
    float q = 0;
    const int a1 = min(fnx, i * N + N - offx), b1 = min(fny, j * N + N - offy), e1 = min(fnz, k * N + N - offz);
    for (int a = max(0, i * N - offx); a < a1; ++a)
        for (int b = max(0, j * N - offy); b < b1; ++b)
            for (int e = max(0, k * N - offz); e < e1; ++e) {
                ++count;
                q += fine[lin3(a, b, e, fnx, fny)];
            }
    coarse[c] = count > 0 ? q / (float)count : 0.f;
}

// out = coarse sampled at the fine sample's world position ((i - fo)*hf), or,
// when fine_in is given, fine_in - that sample (the small-scale difference).
__global__ __launch_bounds__(kBlock) void k_coarse_to_fine(const float* __restrict__ fine_in, float* __restrict__ out,
                                                           int fnx, int fny, int fnz, float fox, float foy, float foz,
                                                           float hf, Grid coarse, double hc) {
    const size_t c = blockIdx.x * (size_t)kBlock + threadIdx.x;
    if (c >= (size_t)fnx * fny * fnz) return;
    const int i = (int)(c % fnx), j = (int)((c / fnx) % fny), k = (int)(c / ((size_t)fnx * fny));
    const float px = ((float)i - fox) * hf, py = ((float)j - foy) * hf, pz = ((float)k - foz) * hf;
    const float s = sample_linear(coarse, hc, px, py, pz);
    out[c] = fine_in ? fine_in[c] - s : s;
}

// Flux matching: the fine faces under a coarse face get the difference between
// the coarse value and their mean. Each coarse cell owns a disjoint fine block.
__global__ __launch_bounds__(kBlock) void k_set_face_boundary(float* __restrict__ fine, int fnx, int fny, int fnz,
                                                              const float* __restrict__ coarse, int cnx, int cny,
                                                              int cnz, int N, int offx, int offy, int offz) {
    const size_t c = blockIdx.x * (size_t)kBlock + threadIdx.x;
    if (c >= (size_t)cnx * cny * cnz) return;
    const int i = (int)(c % cnx), j = (int)((c / cnx) % cny), k = (int)(c / ((size_t)cnx * cny));
    const int a0 = max(0, i * N), b0 = max(0, j * N), e0 = max(0, k * N);
    const int a1 = min(fnx, i * N + offx), b1 = min(fny, j * N + offy), e1 = min(fnz, k * N + offz);
    float q = 0;
    for (int a = a0; a < a1; ++a)
        for (int b = b0; b < b1; ++b)
            for (int e = e0; e < e1; ++e) q += fine[lin3(a, b, e, fnx, fny)];
    const float diff = coarse[c] - q / (float)(N * N);
    for (int a = a0; a < a1; ++a)
        for (int b = b0; b < b1; ++b)
            for (int e = e0; e < e1; ++e) fine[lin3(a, b, e, fnx, fny)] += diff;
}

// Buffer3D bricks: element e of brick (I,J,K) = e >> 9 with brick index
// (K*by + J)*bx + I holds cell (8I + (e & 7), 8J + ((e >> 3) & 7), 8K + (e >> 6)).
// dir 0: linear -> bricks (padding cells set to 0); dir 1: bricks -> linear.
// One thread per brick element, so the brick side reads/writes are contiguous.
__global__ __launch_bounds__(kBlock) void k_bricks(const float* __restrict__ src, float* __restrict__ dst, int nx, int ny,
                                                   int nz, int dir) {
    const int bx = (nx + 7) / 8, by = (ny + 7) / 8, bz = (nz + 7) / 8;
    const size_t e = blockIdx.x * (size_t)kBlock + threadIdx.x;
    if (e >= (size_t)bx * by * bz * 512) return;
    const size_t brick = e >> 9;
    const int w = (int)(e & 511);
    const int I = (int)(brick % bx), J = (int)((brick / bx) % by), K = (int)(brick / ((size_t)bx * by));
    const int i = I * 8 + (w & 7), j = J * 8 + ((w >> 3) & 7), k = K * 8 + (w >> 6);
    const bool in = i < nx && j < ny && k < nz;
    if (dir == 0) {
        dst[e] = in ? src[lin3(i, j, k, nx, ny)] : 0.f;
    } else if (in) {
        dst[lin3(i, j, k, nx, ny)] = src[e];
    }
}

}  // namespace vflip

using namespace vflip;

extern "C" {

int vflip_fine_to_coarse(const float* fine, int fnx, int fny, int fnz, float* coarse, int cnx, int cny, int cnz,
                         int ratio, int offx, int offy, int offz, void* hip_stream) {
    VF_ARG(fine && coarse && fnx > 0 && fny > 0 && fnz > 0 && cnx > 0 && cny > 0 && cnz > 0 && ratio > 0,
           "fine_to_coarse args");
    const size_t n = (size_t)cnx * cny * cnz;
    k_fine_to_coarse<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(fine, fnx, fny, fnz, coarse, cnx, cny, cnz,
                                                                          ratio, offx, offy, offz);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_coarse_to_fine(float* fine, int fnx, int fny, int fnz, float fox, float foy, float foz, float hf,
                         const float* coarse, int cnx, int cny, int cnz, float cox, float coy, float coz, float hc,
                         const float* fine_minus, void* hip_stream) {
    VF_ARG(fine && coarse && fnx > 0 && fny > 0 && fnz > 0 && cnx > 0 && cny > 0 && cnz > 0 && hf > 0 && hc > 0,
           "coarse_to_fine args");
    const size_t n = (size_t)fnx * fny * fnz;
    const Grid g{coarse, cnx, cny, cnz, (double)cox, (double)coy, (double)coz};
    k_coarse_to_fine<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(fine_minus, fine, fnx, fny, fnz, fox, foy,
                                                                          foz, hf, g, (double)hc);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_set_face_boundary(float* fine, int fnx, int fny, int fnz, const float* coarse, int cnx, int cny, int cnz,
                            int ratio, int offx, int offy, int offz, void* hip_stream) {
    VF_ARG(fine && coarse && fnx > 0 && fny > 0 && fnz > 0 && cnx > 0 && cny > 0 && cnz > 0 && ratio > 0,
           "set_face_boundary args");
    VF_ARG(offx <= ratio && offy <= ratio && offz <= ratio && offx >= 0 && offy >= 0 && offz >= 0,
           "face block within one coarse cell (disjoint writes)");
    const size_t n = (size_t)cnx * cny * cnz;
    k_set_face_boundary<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(fine, fnx, fny, fnz, coarse, cnx, cny,
                                                                             cnz, ratio, offx, offy, offz);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_to_bricks(const float* src, int nx, int ny, int nz, float* dst, void* hip_stream) {
    VF_ARG(src && dst && nx > 0 && ny > 0 && nz > 0, "to_bricks args");
    const size_t n = (size_t)((nx + 7) / 8) * ((ny + 7) / 8) * ((nz + 7) / 8) * 512;
    k_bricks<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(src, dst, nx, ny, nz, 0);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_from_bricks(const float* src, int nx, int ny, int nz, float* dst, void* hip_stream) {
    VF_ARG(src && dst && nx > 0 && ny > 0 && nz > 0, "from_bricks args");
    const size_t n = (size_t)((nx + 7) / 8) * ((ny + 7) / 8) * ((nz + 7) / 8) * 512;
    k_bricks<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(src, dst, nx, ny, nz, 1);
    VF_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
