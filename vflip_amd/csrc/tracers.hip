// Smoke tracers: float4 (x, y, z, density) per tracer.
//
// Restates:
//   advect_tracers   S/Smoke_solver3D.h:101-125 (traceRK3 through get_velocity,
//                    push-out where get_solid_phi <= 0, density decay)
//   emit_tracers     S/Smoke_solver3D.h:45-97 (sphere emitter; subdomain tracers
//                    near the solid)
//   write_tracers    S/Smoke_solver3D.h:126-148 (frame record)
// rand() is replaced by the counter hash (frand(a, b) = a + (b - a) U with
// U = (randhash(seed * 0x9E3779B1 + key) >> 8) / 2^24). Compiled without FMA
// contraction so every float expression rounds as the reference's does.
#pragma clang fp contract(off)

#include <cmath>
#include <cstdint>

#include "common.h"
#include "flip_internal.h"
#include "nested_sampling.h"

namespace vflip {

__device__ inline float frand_hash(unsigned seed, unsigned key, float a, float b) {
    const unsigned r = randhash(seed * 0x9E3779B1u + key);
    const float u = (float)(r >> 8) * (1.0f / 16777216.0f);
    return a + (b - a) * u;
}

__global__ __launch_bounds__(kBlock) void k_advect_tracers(int n, float* __restrict__ tr, float dt, float decay,
                                                           vflip_coarse_view C, vflip_subdomain_view S) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const float c1 = (float)(2.0 / 9.0 * dt), c2 = (float)(3.0 / 9.0 * dt);
    const float c3 = (float)((1.0 - c1 - c2) * dt);
    const float t1 = (float)(0.5 * dt), t2 = (float)(0.75 * dt);
    const float4 t = reinterpret_cast<const float4*>(tr)[p];
    float x[3] = {t.x, t.y, t.z};
    float v1[3], v2[3], v3[3], m[3];
    nested_velocity(C, S, x, v1);
#pragma unroll
    for (int a = 0; a < 3; ++a) m[a] = x[a] + t1 * v1[a];
    nested_velocity(C, S, m, v2);
#pragma unroll
    for (int a = 0; a < 3; ++a) m[a] = x[a] + t2 * v2[a];
    nested_velocity(C, S, m, v3);
#pragma unroll
    for (int a = 0; a < 3; ++a) x[a] = x[a] + c1 * v1[a] + c2 * v2[a] + c3 * v3[a];
    const float ph = nested_phi(C, S, x);
    if (ph <= 0) {
        float g[3];
        nested_normal(C, S, x, g);
#pragma unroll
        for (int a = 0; a < 3; ++a) x[a] -= ph * g[a];
    }
    const float d = (float)(t.w / (1.0 + decay * dt));
    reinterpret_cast<float4*>(tr)[p] = make_float4(x[0], x[1], x[2], d);
}

// Sphere emitter candidates t0 .. t0 + k - 1: flag = inside the sphere.
__device__ inline bool sphere_candidate(unsigned seed, unsigned t, float r, float hx, float* xyz) {
    xyz[0] = frand_hash(seed, 3 * t, -r - hx, r + hx);
    xyz[1] = frand_hash(seed, 3 * t + 1, -r - hx, r + hx);
    xyz[2] = frand_hash(seed, 3 * t + 2, -r - hx, r + hx);
    return xyz[0] * xyz[0] + xyz[2] * xyz[2] + xyz[1] * xyz[1] <= r * r;
}

__global__ __launch_bounds__(kBlock) void k_sphere_flag(unsigned t0, int k, unsigned seed, float r, float hx,
                                                        int* __restrict__ flag) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= k) return;
    float xyz[3];
    flag[q] = sphere_candidate(seed, t0 + (unsigned)q, r, hx, xyz);
}

// Accepted candidate q goes to out[base + offset[q]] while that is below n_emit.
__global__ __launch_bounds__(kBlock) void k_sphere_emit(unsigned t0, int k, unsigned seed, float r, float hx,
                                                        float cx, float cy, float cz, const int* __restrict__ offset,
                                                        int base, int n_emit, float* __restrict__ out) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= k) return;
    const int o = offset[q];
    if (offset[q + 1] == o || base + o >= n_emit) return;
    float xyz[3];
    sphere_candidate(seed, t0 + (unsigned)q, r, hx, xyz);
    reinterpret_cast<float4*>(out)[base + o] = make_float4(cx + xyz[0], cy + xyz[1], cz + xyz[2], 1.0f);
}

// Subdomain tracers: 4 candidates per cell, bmin + (i,j,k) h + U h, kept where
// 0 < get_solid_phi < h.
__device__ inline void sub_tracer_pos(const vflip_subdomain_view& S, unsigned seed, size_t c, int q, float* x) {
    const unsigned plane = (unsigned)S.ni * (unsigned)S.nj, cu = (unsigned)c;
    const int k = (int)(cu / plane);
    const int rr = (int)(cu - (unsigned)k * plane);
    const int j = rr / S.ni, i = rr - j * S.ni;
    const int ijk[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u = frand_hash(seed, (unsigned)(3 * (4 * c + (size_t)q) + a), 0.0f, 1.0f);
        x[a] = (S.bmin[a] + (float)ijk[a] * S.h) + u * S.h;
    }
}

__global__ __launch_bounds__(kBlock) void k_sub_tracer_count(size_t ncell, vflip_coarse_view C,
                                                             vflip_subdomain_view S, unsigned seed,
                                                             int* __restrict__ cnt, int* __restrict__ mask) {
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    int m = 0;
    for (int q = 0; q < 4; ++q) {
        float x[3];
        sub_tracer_pos(S, seed, c, q, x);
        const float ph = nested_phi(C, S, x);
        if (ph > 0 && ph < S.h) m |= 1 << q;
    }
    cnt[c] = __popc(m);
    mask[c] = m;
}

__global__ __launch_bounds__(kBlock) void k_sub_tracer_emit(size_t ncell, vflip_subdomain_view S, unsigned seed,
                                                            const int* __restrict__ offset,
                                                            const int* __restrict__ mask, float* __restrict__ out) {
    const size_t c = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    const int m = mask[c];
    size_t o = (size_t)offset[c];
    for (int q = 0; q < 4; ++q) {
        if (!((m >> q) & 1)) continue;
        float x[3];
        sub_tracer_pos(S, seed, c, q, x);
        reinterpret_cast<float4*>(out)[o++] = make_float4(x[0], x[1], x[2], 1.0f);
    }
}

__global__ __launch_bounds__(kBlock) void k_tracer_frame(int n, const float* __restrict__ tr, float lx,
                                                         float* __restrict__ out) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const float4 t = reinterpret_cast<const float4*>(tr)[p];
    reinterpret_cast<float4*>(out)[p] =
        make_float4((float)(t.x / lx - 0.5), t.y / lx, (float)(t.z / lx - 0.5), t.w);
}

static bool tview_ok(const vflip_coarse_view* C) {
    return C && C->ni > 0 && C->nj > 0 && C->nk > 0 && C->h > 0 && C->u && C->v && C->w && C->solid_phi;
}
static bool tview_ok(const vflip_subdomain_view* S) {
    return S && S->ni > 0 && S->nj > 0 && S->nk > 0 && S->h > 0 && S->u && S->v && S->w && S->solid_phi;
}


}  // namespace vflip

using namespace vflip;

int vflip_advect_tracers(const vflip_coarse_view* C, const vflip_subdomain_view* S, float dt, float dens_decay,
                         int n, float* tracers, void* hip_stream) {
    VF_ARG(tview_ok(C) && (!S || tview_ok(S)) && n >= 0, "advect tracers args");
    VF_ARG(n == 0 || tracers, "advect tracers array");
    VF_ARG(((uintptr_t)tracers & 15) == 0, "tracers must be 16-byte aligned (float4)");
    const vflip_subdomain_view sv = S ? *S : no_subdomain();
    if (n > 0)
        k_advect_tracers<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(n, tracers, dt, dens_decay, *C, sv);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_emit_tracers_sphere(vflip_particles* P, unsigned seed, int n_emit, float cx, float cy, float cz, float r,
                              float hx, float* out) {
    VF_ARG(P && n_emit >= 0 && r > 0 && hx >= 0, "sphere emitter args");
    VF_ARG(n_emit == 0 || out, "sphere emitter output");
    VF_ARG(((uintptr_t)out & 15) == 0, "output must be 16-byte aligned (float4)");
    hipStream_t s = P->stream;
    const int k = 2 * n_emit + 4096;  // candidates per round (about half are accepted)
    if (int rc = ensure_selection_scratch(P, (size_t)k)) return rc;
    int got = 0;
    unsigned t0 = 0;
    for (int round = 0; got < n_emit; ++round) {
        VF_ARG(round < 64, "sphere emitter: acceptance too low (r against hx)");
        k_sphere_flag<<<grid_for(k), kBlock, 0, s>>>(t0, k, seed, r, hx, P->pflag);
        scan_exclusive_i32(P->pflag, (size_t)k, P->ppartial, s);
        k_sphere_emit<<<grid_for(k), kBlock, 0, s>>>(t0, k, seed, r, hx, cx, cy, cz, P->pflag, got, n_emit, out);
        VF_CHECK_LAUNCH();
        int acc = 0;
        VF_HIP(hipMemcpyAsync(&acc, P->pflag + k, sizeof(int), hipMemcpyDeviceToHost, s));
        VF_HIP(hipStreamSynchronize(s));
        got += acc;
        t0 += (unsigned)k;
    }
    return 0;
}

int vflip_emit_tracers_sub(vflip_particles* P, const vflip_coarse_view* C, const vflip_subdomain_view* S,
                           unsigned seed, int cap, float* out, int* nout) {
    VF_ARG(P && tview_ok(C) && tview_ok(S) && nout && cap >= 0, "subdomain tracer args");
    VF_ARG(P->nx == S->ni && P->ny == S->nj && P->nz == S->nk, "subdomain tracers: P must be the subdomain's workspace");
    VF_ARG(cap == 0 || out, "subdomain tracer output");
    VF_ARG(((uintptr_t)out & 15) == 0, "output must be 16-byte aligned (float4)");
    hipStream_t s = P->stream;
    const size_t nc = P->ncell;
    if (!P->reo && hipMalloc(&P->reo, (2 * nc + 1) * sizeof(int)) != hipSuccess) {
        P->reo = nullptr;
        set_error("tracer workspace of %zu cells: out of device memory", nc);
        return VFLIP_ERR_NOMEM;
    }
    int* cnt = P->reo;
    int* mask = P->reo + nc + 1;
    k_sub_tracer_count<<<grid_for(nc), kBlock, 0, s>>>(nc, *C, *S, seed, cnt, mask);
    scan_exclusive_i32(cnt, nc, P->partial, s);
    VF_CHECK_LAUNCH();
    int total = 0;
    VF_HIP(hipMemcpyAsync(&total, cnt + nc, sizeof(int), hipMemcpyDeviceToHost, s));
    VF_HIP(hipStreamSynchronize(s));
    *nout = total;
    if (total > cap) {
        set_error("subdomain tracers: %d do not fit the output capacity %d", total, cap);
        return VFLIP_ERR_CAPACITY;
    }
    if (total > 0) k_sub_tracer_emit<<<grid_for(nc), kBlock, 0, s>>>(nc, *S, seed, cnt, mask, out);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_tracer_frame(int n, const float* tracers, float lx, float* out, void* hip_stream) {
    VF_ARG(n >= 0 && (n == 0 || (tracers && out)), "tracer frame args");
    VF_ARG(((uintptr_t)tracers & 15) == 0 && ((uintptr_t)out & 15) == 0, "tracer arrays must be 16-byte aligned");
    if (n > 0) k_tracer_frame<<<grid_for(n), kBlock, 0, (hipStream_t)hip_stream>>>(n, tracers, lx, out);
    VF_CHECK_LAUNCH();
    return 0;
}
