// Machine probes behind the roofline statements of the solver's kernels (a
// separate library, libvflip_probe.so; not part of the solver's C ABI):
//   * streaming bandwidth of plain passes shaped like the solver's kernels: a
//     copy (1 read, 1 write), the smoother's pattern (6 reads, 1 write) and a
//     read-only sum, float4 per lane, over arrays of a given size;
//   * the cost of a grid-wide barrier (cooperative launch) against the number
//     of workgroups, and of back-to-back small kernel launches.
// Every kernel stays inside the arrays it is given (grid-stride loops bounded by
// the element count); the barrier kernel touches one word per workgroup.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <chrono>

namespace {

__global__ __launch_bounds__(256) void k_copy4(const float4* __restrict__ a, float4* __restrict__ o, size_t n4) {
    for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n4; e += (size_t)gridDim.x * 256) o[e] = a[e];
}

// o = a + b + c + d + e + f (the level-0 sweep's six input streams and one output)
__global__ __launch_bounds__(256) void k_six_to_one(const float4* __restrict__ a, const float4* __restrict__ b,
                                                    const float4* __restrict__ c, const float4* __restrict__ d,
                                                    const float4* __restrict__ e, const float4* __restrict__ f,
                                                    float4* __restrict__ o, size_t n4) {
    for (size_t q = blockIdx.x * (size_t)256 + threadIdx.x; q < n4; q += (size_t)gridDim.x * 256) {
        const float4 x = a[q], y = b[q], z = c[q], u = d[q], v = e[q], w = f[q];
        o[q] = make_float4(x.x + y.x + z.x + u.x + v.x + w.x, x.y + y.y + z.y + u.y + v.y + w.y,
                           x.z + y.z + z.z + u.z + v.z + w.z, x.w + y.w + z.w + u.w + v.w + w.w);
    }
}

// per-workgroup partial sums of a (read-only stream); partial has gridDim.x entries
__global__ __launch_bounds__(256) void k_read_sum(const float4* __restrict__ a, size_t n4, float* __restrict__ partial) {
    __shared__ float red[256];
    float s = 0.f;
    for (size_t q = blockIdx.x * (size_t)256 + threadIdx.x; q < n4; q += (size_t)gridDim.x * 256) {
        const float4 x = a[q];
        s += (x.x + x.y) + (x.z + x.w);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// The mixed-precision sweep's traffic shape, 20 B per element, two layouts:
// seven streams (b, x in, four 2-byte coefficient arrays, x out) or four
// (b, x in, one 8-byte coefficient record per element, x out). Each thread
// takes 4 consecutive elements.
__global__ __launch_bounds__(256) void k_sweep7(const float4* __restrict__ b, const float4* __restrict__ x,
                                                const uint2* __restrict__ d, const uint2* __restrict__ ax,
                                                const uint2* __restrict__ ay, const uint2* __restrict__ az,
                                                float4* __restrict__ o, size_t n4) {
    for (size_t q = blockIdx.x * (size_t)256 + threadIdx.x; q < n4; q += (size_t)gridDim.x * 256) {
        const float4 bb = b[q], xx = x[q];
        const uint2 c0 = d[q], c1 = ax[q], c2 = ay[q], c3 = az[q];
        const float s = __uint_as_float(c0.x ^ c1.x ^ c2.x ^ c3.x) + __uint_as_float(c0.y ^ c1.y ^ c2.y ^ c3.y);
        o[q] = make_float4(bb.x + xx.x + s, bb.y + xx.y, bb.z + xx.z, bb.w + xx.w);
    }
}
__global__ __launch_bounds__(256) void k_sweep4(const float4* __restrict__ b, const float4* __restrict__ x,
                                                const uint4* __restrict__ coef, float4* __restrict__ o, size_t n4) {
    for (size_t q = blockIdx.x * (size_t)256 + threadIdx.x; q < n4; q += (size_t)gridDim.x * 256) {
        const float4 bb = b[q], xx = x[q];
        const uint4 c0 = coef[2 * q], c1 = coef[2 * q + 1];
        const float s = __uint_as_float(c0.x ^ c0.y ^ c0.z ^ c0.w) + __uint_as_float(c1.x ^ c1.y ^ c1.z ^ c1.w);
        o[q] = make_float4(bb.x + xx.x + s, bb.y + xx.y, bb.z + xx.z, bb.w + xx.w);
    }
}

__global__ __launch_bounds__(256) void k_barriers(float* __restrict__ slot, int nbar) {
    cooperative_groups::grid_group g = cooperative_groups::this_grid();
    for (int i = 0; i < nbar; ++i) {
        if (threadIdx.x == 0) slot[blockIdx.x] += 1.f;
        g.sync();
    }
}

__global__ void k_tiny(float* __restrict__ slot) {
    if (threadIdx.x == 0) slot[blockIdx.x] += 1.f;
}

int fail(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }
#define PR_CHECK(x)                          \
    do {                                     \
        const hipError_t e_ = (x);           \
        if (e_ != hipSuccess) return fail(e_); \
    } while (0)

}  // namespace

extern "C" {

// Average milliseconds of one pass over n floats per array (n a multiple of 4),
// kind 0: copy, 1: six reads and one write, 2: read-only sum, 3: the
// mixed-precision sweep's 20 B per element as seven streams (k_sweep7: bufs[2..5]
// are read as n/2 floats each), 4: the same bytes as four streams (k_sweep4:
// bufs[2] and bufs[3] hold the n 8-byte records). bufs: 7 device arrays of n
// floats (the copy uses bufs[0] -> bufs[6]; the sum bufs[0] and bufs[6] as
// partials). Returns 0 or a negative HIP error.
int vflip_probe_stream(int kind, float* const* bufs, size_t n, int blocks, int reps, double* ms_per_pass) {
    if (!bufs || !ms_per_pass || n % 4 || blocks <= 0 || reps <= 0 || kind < 0 || kind > 4) return -1;
    if (kind == 4 && bufs[3] != bufs[2] + n) return -1;  // the records span bufs[2] and bufs[3]
    const size_t n4 = n / 4;
    const float4* in[6];
    for (int q = 0; q < 6; ++q) in[q] = reinterpret_cast<const float4*>(bufs[q]);
    float4* out = reinterpret_cast<float4*>(bufs[6]);
    if (kind == 2 && (size_t)blocks > n) return -1;  // partials live in bufs[6]
    auto launch = [&]() {
        if (kind == 0) k_copy4<<<blocks, 256>>>(in[0], out, n4);
        else if (kind == 1) k_six_to_one<<<blocks, 256>>>(in[0], in[1], in[2], in[3], in[4], in[5], out, n4);
        else if (kind == 2) k_read_sum<<<blocks, 256>>>(in[0], n4, bufs[6]);
        else if (kind == 3)
            k_sweep7<<<blocks, 256>>>(in[0], in[1], reinterpret_cast<const uint2*>(bufs[2]),
                                      reinterpret_cast<const uint2*>(bufs[3]), reinterpret_cast<const uint2*>(bufs[4]),
                                      reinterpret_cast<const uint2*>(bufs[5]), out, n4);
        else k_sweep4<<<blocks, 256>>>(in[0], in[1], reinterpret_cast<const uint4*>(bufs[2]), out, n4);
    };
    hipEvent_t a, b;
    PR_CHECK(hipEventCreate(&a));
    PR_CHECK(hipEventCreate(&b));
    launch();
    launch();
    PR_CHECK(hipEventRecord(a, nullptr));
    for (int r = 0; r < reps; ++r) launch();
    PR_CHECK(hipEventRecord(b, nullptr));
    PR_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    PR_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_pass = ms / reps;
    return 0;
}

// Microseconds per grid-wide barrier of a cooperative launch with `blocks`
// workgroups of 256 threads (nbar barriers timed); slot: >= blocks floats.
int vflip_probe_grid_barrier(float* slot, int blocks, int nbar, double* us_per_barrier) {
    if (!slot || !us_per_barrier || blocks <= 0 || nbar <= 0) return -1;
    void* args[] = {&slot, &nbar};
    PR_CHECK(hipLaunchCooperativeKernel((const void*)k_barriers, dim3(blocks), dim3(256), args, 0, nullptr));
    PR_CHECK(hipDeviceSynchronize());
    hipEvent_t a, b;
    PR_CHECK(hipEventCreate(&a));
    PR_CHECK(hipEventCreate(&b));
    PR_CHECK(hipEventRecord(a, nullptr));
    PR_CHECK(hipLaunchCooperativeKernel((const void*)k_barriers, dim3(blocks), dim3(256), args, 0, nullptr));
    PR_CHECK(hipEventRecord(b, nullptr));
    PR_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    PR_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *us_per_barrier = ms * 1e3 / nbar;
    return 0;
}

// Back-to-back launches of a one-word kernel with `blocks` workgroups: GPU
// microseconds per launch and host microseconds per enqueue; slot: >= blocks floats.
int vflip_probe_launch(float* slot, int blocks, int reps, double* gpu_us, double* host_us) {
    if (!slot || !gpu_us || !host_us || blocks <= 0 || reps <= 0) return -1;
    for (int r = 0; r < 10; ++r) k_tiny<<<blocks, 256>>>(slot);
    PR_CHECK(hipDeviceSynchronize());
    hipEvent_t a, b;
    PR_CHECK(hipEventCreate(&a));
    PR_CHECK(hipEventCreate(&b));
    const auto t0 = std::chrono::steady_clock::now();
    PR_CHECK(hipEventRecord(a, nullptr));
    for (int r = 0; r < reps; ++r) k_tiny<<<blocks, 256>>>(slot);
    PR_CHECK(hipEventRecord(b, nullptr));
    const auto t1 = std::chrono::steady_clock::now();
    PR_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    PR_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *gpu_us = ms * 1e3 / reps;
    *host_us = std::chrono::duration<double, std::micro>(t1 - t0).count() / reps;
    return 0;
}

}  // extern "C"
