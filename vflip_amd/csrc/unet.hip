// Learned warm start for the pressure solve: a 3D U-Net (4 levels, 16/32/64/128
// channels, 3x3x3 convolutions, BF16 weights and activations, FP32
// accumulation) that maps (divergence, solid mask, liquid mask) to pressure.
// Forward, backward and Adam run on the GPU through the kernels of
// unet_kernels.h (matrix-core convolutions for the forward, backward-data and
// weight-gradient passes). The solver keeps the last word: the prediction only
// seeds MGPCG, which runs to the same tolerance.
//
// Network (level l has grid n / 2^l):
//   enc l: conv(Cin -> C_l), conv(C_l -> C_l)          C = 16, 32, 64, 128
//   between levels: 2x2x2 average pooling
//   dec l: conv([up(dec l+1) | enc l] -> C_l), conv(C_l -> C_l)   (nearest upsampling, concat)
//   out:   conv(16 -> 1), no activation
// There is no counterpart in the reference; the interface mirrors the solver's
// warm_start option (the PCG takes the prediction as its initial guess).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.h"
#include "unet_kernels.h"

using namespace vflip;
using namespace vflip::unet;

namespace {

constexpr int kLayers = 15;
constexpr int kTensors = kLayers + 4;  // layer outputs, pooled P0..P2, network input
constexpr int kIn = kLayers + 3;
constexpr int kP0 = kLayers, kP1 = kLayers + 1, kP2 = kLayers + 2;

struct In {
    int t, C, up;
};
struct LayerDef {
    In in[2];
    int nin;
    int cout, relu, level, cin_real;
};

// clang-format off
const LayerDef kNet[kLayers] = {
    {{{kIn, 16, 0}}, 1, 16, 1, 0, 3},     // 0  enc0a
    {{{0, 16, 0}}, 1, 16, 1, 0, 16},      // 1  enc0b (skip 0)
    {{{kP0, 16, 0}}, 1, 32, 1, 1, 16},    // 2  enc1a
    {{{2, 32, 0}}, 1, 32, 1, 1, 32},      // 3  enc1b (skip 1)
    {{{kP1, 32, 0}}, 1, 64, 1, 2, 32},    // 4  enc2a
    {{{4, 64, 0}}, 1, 64, 1, 2, 64},      // 5  enc2b (skip 2)
    {{{kP2, 64, 0}}, 1, 128, 1, 3, 64},   // 6  bottleneck a
    {{{6, 128, 0}}, 1, 128, 1, 3, 128},   // 7  bottleneck b
    {{{7, 128, 1}, {5, 64, 0}}, 2, 64, 1, 2, 192},  // 8  dec2a
    {{{8, 64, 0}}, 1, 64, 1, 2, 64},      // 9  dec2b
    {{{9, 64, 1}, {3, 32, 0}}, 2, 32, 1, 1, 96},    // 10 dec1a
    {{{10, 32, 0}}, 1, 32, 1, 1, 32},     // 11 dec1b
    {{{11, 32, 1}, {1, 16, 0}}, 2, 16, 1, 0, 48},   // 12 dec0a
    {{{12, 16, 0}}, 1, 16, 1, 0, 16},     // 13 dec0b
    {{{13, 16, 0}}, 1, 16, 0, 0, 16},     // 14 out (channel 0 is the prediction)
};
// clang-format on
// pooled tensors: P0 = pool(layer 1), P1 = pool(layer 3), P2 = pool(layer 5)
const int kPoolOf[3] = {1, 3, 5};

int cin_of(const LayerDef& d) { return d.in[0].C + (d.nin > 1 ? d.in[1].C : 0); }
int ck_of(const LayerDef& d) {
    bool all32 = d.in[0].C % 32 == 0 && (d.nin < 2 || d.in[1].C % 32 == 0);
    return all32 ? 32 : 16;
}
int kw_of(const LayerDef& d) { return (ck_of(d) == 32 ? 27 : 28) * cin_of(d); }
int ckb_of(const LayerDef& d) { return d.cout % 32 == 0 ? 32 : 16; }
int kwb_of(const LayerDef& d) { return (ckb_of(d) == 32 ? 27 : 28) * d.cout; }

unsigned grid256(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 16384); }

__global__ void k_bf16_to_f32(const bf16* __restrict__ a, size_t n, float* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)a[i];
}

}  // namespace

struct vflip_unet {
    hipStream_t stream = nullptr;
    // parameters: [weights of all layers][biases of all layers], fp32 master
    size_t nw = 0, nb = 0;
    size_t woff[kLayers], boff[kLayers], wboff[kLayers];
    float* params = nullptr;
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    bf16* w16 = nullptr;   // forward weights (same offsets as the master weights)
    bf16* wb16 = nullptr;  // backward-data weights
    size_t nwb = 0;
    // per-grid buffers
    int nx = 0, ny = 0, nz = 0;
    bf16* act[kTensors] = {};
    float* grad[kTensors] = {};
    float* pred = nullptr;  // [V0]
    bf16* gpre = nullptr;
    float* gx = nullptr;
    float* gpred = nullptr;
    double* loss = nullptr;
    double* eacc = nullptr;  // device: e'Ae, x'Ax of the energy loss
    float* scale = nullptr;  // device: max|rhs| of the current input (automatic normalisation)
    float gain = 1.0f;       // pressure unit of the automatic normalisation: max|rhs| * gain
    std::vector<void*> grid_allocs;
    float* splitk = nullptr;  // FP32 partials of the split convolutions (per grid)
    // warm-start workspace on the network grid
    int wx = 0, wy = 0, wz = 0;
    float *w_rhs = nullptr, *w_solid = nullptr, *w_liquid = nullptr, *w_pred = nullptr;
};

namespace {

int level_dims(const vflip_unet* N, int level, int& x, int& y, int& z) {
    x = N->nx >> level;
    y = N->ny >> level;
    z = N->nz >> level;
    return 0;
}
size_t level_vox(const vflip_unet* N, int level) {
    int x, y, z;
    level_dims(N, level, x, y, z);
    return (size_t)x * y * z;
}
int tensor_level(int t) {
    if (t == kIn) return 0;
    if (t >= kP0 && t <= kP2) return t - kP0 + 1;
    return kNet[t].level;
}
int tensor_channels(int t) {
    if (t == kIn) return 16;
    if (t >= kP0 && t <= kP2) return kNet[kPoolOf[t - kP0]].cout;
    return kNet[t].cout;
}

void free_grid(vflip_unet* N) {
    for (void* p : N->grid_allocs) (void)hipFree(p);
    N->grid_allocs.clear();
    for (int t = 0; t < kTensors; ++t) {
        N->act[t] = nullptr;
        N->grad[t] = nullptr;
    }
    N->pred = N->gx = N->gpred = nullptr;
    N->gpre = nullptr;
    N->splitk = nullptr;
    N->nx = N->ny = N->nz = 0;
}

template <typename T>
int galloc(vflip_unet* N, T** p, size_t n) {
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        set_error("U-Net buffers: out of device memory");
        return VFLIP_ERR_NOMEM;
    }
    N->grid_allocs.push_back(*p);
    return 0;
}

// Launch plan of a convolution: output channels split into g groups until the
// layer has 256 workgroups and each workgroup at most 6 16-channel blocks (or
// one block per workgroup); then, for the small grids, the K loop into ks
// slices (powers of two) until about 256 workgroups, each slice at least 9 of
// the K steps. Groups come first: a split K loop adds an FP32 partial per
// slice and a reduction launch.
int conv_plan(int ntiles, int NBall, int cin, int ck, int& g) {
    g = 1;
    while ((ntiles * g < 256 || NBall / g > 6) && NBall % (2 * g) == 0) g *= 2;
    const int total = ((cin + ck - 1) / ck) * (ck == 32 ? 27 : 15);
    int ks = 1;
    while (ntiles * g * ks * 2 <= 256 && total / (ks * 2) >= 9) ks *= 2;
    return ks;
}

int level_tiles(const vflip_unet* N, int level) {
    int x, y, z;
    level_dims(N, level, x, y, z);
    return ((x + kTX - 1) / kTX) * ((y + kTY - 1) / kTY) * ((z + kTZ - 1) / kTZ);
}

// Largest partial buffer of the forward and backward-data convolutions.
size_t splitk_floats(const vflip_unet* N) {
    size_t need = 0;
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = kNet[l];
        const size_t V = level_vox(N, d.level);
        const int nt = level_tiles(N, d.level);
        int g;
        int ks = conv_plan(nt, d.cout / 16, cin_of(d), ck_of(d), g);
        if (ks > 1) need = std::max(need, (size_t)ks * V * d.cout);
        if (l == 0) continue;
        ks = conv_plan(nt, cin_of(d) / 16, d.cout, ckb_of(d), g);
        if (ks > 1) need = std::max(need, (size_t)ks * V * cin_of(d));
    }
    return need;
}

int ensure_grid(vflip_unet* N, int nx, int ny, int nz) {
    if (N->nx == nx && N->ny == ny && N->nz == nz) return 0;
    free_grid(N);
    N->nx = nx;
    N->ny = ny;
    N->nz = nz;
    size_t gmax = 0, gxmax = 0;
    for (int t = 0; t < kTensors; ++t) {
        if (t == kLayers - 1) continue;  // the output layer writes pred
        const size_t n = level_vox(N, tensor_level(t)) * tensor_channels(t);
        if (int rc = galloc(N, &N->act[t], n)) return rc;
        if (t != kIn)
            if (int rc = galloc(N, &N->grad[t], n)) return rc;
    }
    for (int l = 0; l < kLayers; ++l) {
        const size_t V = level_vox(N, kNet[l].level);
        gmax = std::max(gmax, V * kNet[l].cout);
        gxmax = std::max(gxmax, V * cin_of(kNet[l]));
    }
    if (int rc = galloc(N, &N->gpre, gmax)) return rc;
    if (int rc = galloc(N, &N->gx, gxmax)) return rc;
    if (int rc = galloc(N, &N->pred, level_vox(N, 0))) return rc;
    if (int rc = galloc(N, &N->gpred, level_vox(N, 0))) return rc;
    if (int rc = galloc(N, &N->loss, 1)) return rc;
    if (const size_t nk = splitk_floats(N))
        if (int rc = galloc(N, &N->splitk, nk)) return rc;
    return 0;
}

ConvSrc make_src(const vflip_unet* N, const In& in) { return ConvSrc{N->act[in.t], in.C, in.up}; }

// Workgroups of k_conv3_tiles that fit on the current device at once (0 if the
// query fails), cached per device.
int tiles_grid() {
    static int cache[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    if (cache[dev] == 0) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_conv3_tiles, kConvThreads, 0) != hipSuccess)
            return 0;
        cache[dev] = cus * per;
    }
    return cache[dev];
}

// One convolution by conv_plan: ntiles x groups x ksplit workgroups, and for a
// split one the reduction of its partials (workspace sized by splitk_floats).
// Unsplit 16-channel layers into 16 outputs on grids of at least four rounds of
// workgroups run persistent (k_conv3_tiles).
// Returns whether the kernels also wrote a.pool (otherwise the caller pools).
template <int CK>
bool launch_conv(ConvArgs a, int NBall, hipStream_t s, float* ws) {
    int g;
    const int ks = conv_plan(a.ntiles, NBall, a.cin, CK, g);
    a.groups = g;
    a.ksplit = ks;
    a.part = ks > 1 ? ws : nullptr;
    const int NB = NBall / g;
    // (a split layer pools in its reduction when that still has 256 workgroups of
    // 8-voxel sums; on smaller grids the separate passes are faster)
    const bool pool_req = a.pool && a.out && !a.out32;
    const size_t coarse = (size_t)(a.nx / 2) * (a.ny / 2) * (a.nz / 2) * a.cout;
    const bool pooled = pool_req && (ks > 1 ? (a.nx | a.ny | a.nz) % 2 == 0 && coarse >= 256 * 256
                                            : conv_tile_in_lds(CK, NB, false));
    if (!pooled) a.pool = nullptr;
    if (CK == 16 && ks == 1 && g == 1 && NB == 1) {
        const bool bf16_out = a.out && !a.out32;
        const bool ch0_out = !a.out && a.out32 && a.out32_ch0 && (a.nx & 3) == 0;
        const int pg = tiles_grid();
        if ((bf16_out || ch0_out) && pg > 0 && a.ntiles >= 4 * pg) {
            k_conv3_tiles<<<pg, kConvThreads, 0, s>>>(a);
            return pooled;
        }
    }
    const dim3 grid(a.ntiles * g * ks), block(kConvThreads);
    if (ks > 1) {
        switch (NB) {
            case 1: k_conv3<CK, 1, true><<<grid, block, 0, s>>>(a); break;
            case 2: k_conv3<CK, 2, true><<<grid, block, 0, s>>>(a); break;
            case 3: k_conv3<CK, 3, true><<<grid, block, 0, s>>>(a); break;
            case 4: k_conv3<CK, 4, true><<<grid, block, 0, s>>>(a); break;
            case 6: k_conv3<CK, 6, true><<<grid, block, 0, s>>>(a); break;
            case 8: k_conv3<CK, 8, true><<<grid, block, 0, s>>>(a); break;
            case 12: k_conv3<CK, 12, true><<<grid, block, 0, s>>>(a); break;
            default: break;
        }
    } else {
        switch (NB) {
            case 1: k_conv3<CK, 1, false><<<grid, block, 0, s>>>(a); break;
            case 2: k_conv3<CK, 2, false><<<grid, block, 0, s>>>(a); break;
            case 3: k_conv3<CK, 3, false><<<grid, block, 0, s>>>(a); break;
            case 4: k_conv3<CK, 4, false><<<grid, block, 0, s>>>(a); break;
            case 6: k_conv3<CK, 6, false><<<grid, block, 0, s>>>(a); break;
            case 8: k_conv3<CK, 8, false><<<grid, block, 0, s>>>(a); break;
            case 12: k_conv3<CK, 12, false><<<grid, block, 0, s>>>(a); break;
            default: break;
        }
    }
    if (ks > 1) {
        const size_t n = (size_t)a.nx * a.ny * a.nz * a.cout;
        if (pooled) k_conv_splitk_reduce_pool<<<grid256(n / 8), 256, 0, s>>>(a);
        else k_conv_splitk_reduce<<<grid256(n), 256, 0, s>>>(a);
    }
    return pooled;
}

bool conv(const vflip_unet* N, const ConvArgs& a, int ck, hipStream_t s) {
    if (ck == 32) return launch_conv<32>(a, a.cout / 16, s, N->splitk);
    return launch_conv<16>(a, a.cout / 16, s, N->splitk);
}

ConvArgs base_args(const vflip_unet* N, int level) {
    ConvArgs a{};
    level_dims(N, level, a.nx, a.ny, a.nz);
    a.tiles_x = (a.nx + kTX - 1) / kTX;
    a.tiles_y = (a.ny + kTY - 1) / kTY;
    a.ntiles = a.tiles_x * a.tiles_y * ((a.nz + kTZ - 1) / kTZ);
    return a;
}

void forward_layers(vflip_unet* N) {
    hipStream_t s = N->stream;
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = kNet[l];
        // layers 1, 3, 5 feed the pooled input of the level below
        const int pool_t = l == 1 ? kP0 : l == 3 ? kP1 : l == 5 ? kP2 : -1;
        ConvArgs a = base_args(N, d.level);
        a.nsrc = d.nin;
        for (int q = 0; q < d.nin; ++q) a.src[q] = make_src(N, d.in[q]);
        a.cin = cin_of(d);
        a.cout = d.cout;
        a.kw = kw_of(d);
        a.w = N->w16 + N->woff[l];
        a.bias = N->params + N->nw + N->boff[l];
        a.relu = d.relu;
        if (l == kLayers - 1) {
            a.out32 = N->pred;
            a.out32_ch0 = 1;
        } else {
            a.out = N->act[l];
        }
        if (pool_t >= 0) a.pool = N->act[pool_t];
        if (!conv(N, a, ck_of(d), s) && pool_t >= 0) {
            int x, y, z;
            level_dims(N, d.level, x, y, z);
            const size_t n = level_vox(N, d.level + 1) * d.cout;
            k_avgpool<<<grid256(n), 256, 0, s>>>(N->act[l], x, y, z, d.cout, N->act[pool_t]);
        }
    }
}

template <int NCB>
void launch_wgrad(const WgradArgs& w, int cin, size_t V, hipStream_t s) {
    const dim3 grid(27 * (cin / 16), (unsigned)((V + w.chunk - 1) / w.chunk));
    k_conv3_wgrad<NCB><<<grid, 256, 0, s>>>(w);
}

void backward_layers(vflip_unet* N) {
    hipStream_t s = N->stream;
    for (int t = 0; t < kTensors; ++t)
        if (N->grad[t]) {
            const size_t n = level_vox(N, tensor_level(t)) * tensor_channels(t);
            (void)hipMemsetAsync(N->grad[t], 0, n * sizeof(float), s);
        }
    for (int l = kLayers - 1; l >= 0; --l) {
        const LayerDef& d = kNet[l];
        const size_t V = level_vox(N, d.level);
        // gradient at the convolution output, ReLU mask applied, bf16
        if (l == kLayers - 1) k_grad_ch0<<<grid256(V), 256, 0, s>>>(N->gpred, V, N->gpre);
        else k_relu_bwd<<<grid256(V * d.cout), 256, 0, s>>>(N->grad[l], N->act[l], V * d.cout, d.relu, N->gpre);
        // weight and bias gradients
        WgradArgs w{};
        w.nsrc = d.nin;
        for (int q = 0; q < d.nin; ++q) w.src[q] = make_src(N, d.in[q]);
        level_dims(N, d.level, w.nx, w.ny, w.nz);
        w.cin = cin_of(d);
        w.cout = d.cout;
        w.kw = kw_of(d);
        w.g = N->gpre;
        w.dw = N->grads + N->woff[l];
        w.chunk = 4096;
        switch (d.cout / 16) {
            case 1: launch_wgrad<1>(w, w.cin, V, s); break;
            case 2: launch_wgrad<2>(w, w.cin, V, s); break;
            case 4: launch_wgrad<4>(w, w.cin, V, s); break;
            case 8: launch_wgrad<8>(w, w.cin, V, s); break;
        }
        k_bias_grad<<<(unsigned)std::min<size_t>((V + 255) / 256, 512), 256, 0, s>>>(N->gpre, V, d.cout,
                                                                                       N->grads + N->nw + N->boff[l]);
        if (l == 0) break;
        // gradient with respect to the layer input: transposed convolution
        ConvArgs a = base_args(N, d.level);
        a.nsrc = 1;
        a.src[0] = ConvSrc{N->gpre, d.cout, 0};
        a.cin = d.cout;
        a.cout = cin_of(d);
        a.kw = kwb_of(d);
        a.w = N->wb16 + N->wboff[l];
        a.out32 = N->gx;
        conv(N, a, ckb_of(d), s);
        if (d.nin == 1) {
            const In& in = d.in[0];
            const size_t n = V * in.C;
            k_concat_bwd_skip<<<grid256(n), 256, 0, s>>>(N->gx, V, 0, in.C, N->grad[in.t]);
            if (in.t >= kP0 && in.t <= kP2) {  // through the pooling into the level above
                const int src = kPoolOf[in.t - kP0];
                int x, y, z;
                level_dims(N, kNet[src].level, x, y, z);
                const size_t nf = level_vox(N, kNet[src].level) * kNet[src].cout;
                k_avgpool_bwd<<<grid256(nf), 256, 0, s>>>(N->grad[in.t], x, y, z, kNet[src].cout, N->grad[src]);
            }
        } else {
            const In &up = d.in[0], &skip = d.in[1];
            int x, y, z;
            level_dims(N, d.level, x, y, z);
            const size_t nc = level_vox(N, d.level + 1) * up.C;
            k_concat_bwd_up<<<grid256(nc), 256, 0, s>>>(N->gx, x, y, z, up.C, skip.C, N->grad[up.t]);
            k_concat_bwd_skip<<<grid256(V * skip.C), 256, 0, s>>>(N->gx, V, up.C, skip.C, N->grad[skip.t]);
        }
    }
}

void refresh_bwd_weights(vflip_unet* N) {
    for (int l = 1; l < kLayers; ++l) {
        const LayerDef& d = kNet[l];
        const size_t n = (size_t)cin_of(d) * kwb_of(d);
        k_flip_weights<<<grid256(n), 256, 0, N->stream>>>(N->params + N->woff[l], cin_of(d), d.cout, kw_of(d),
                                                          kwb_of(d), N->wb16 + N->wboff[l]);
    }
}

// splitmix64 -> uniform -> Box-Muller normal, for reproducible initialisation.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

bool dims_ok(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0 && nx % 8 == 0 && ny % 8 == 0 && nz % 8 == 0; }

// in_scale > 0: the input is rhs * in_scale and the output the raw network
// value; in_scale == 0: automatic, s = max|rhs| on the device, input rhs / s,
// output (and training target) in units of s. Returns the device scale or null.
const float* prepare_scale(vflip_unet* N, const float* rhs, size_t V, float in_scale) {
    if (in_scale > 0.f) return nullptr;
    k_zero1<<<1, 1, 0, N->stream>>>(N->scale);
    k_absmax_to<<<grid256(V), 256, 0, N->stream>>>(rhs, V, N->scale);
    return N->scale;
}

int ensure_warm_grid(vflip_unet* N, int cx, int cy, int cz) {
    if (N->wx == cx && N->wy == cy && N->wz == cz) return 0;
    for (float* p : {N->w_rhs, N->w_solid, N->w_liquid, N->w_pred})
        if (p) (void)hipFree(p);
    N->w_rhs = N->w_solid = N->w_liquid = N->w_pred = nullptr;
    N->wx = N->wy = N->wz = 0;
    const size_t n = (size_t)cx * cy * cz;
    if (hipMalloc(&N->w_rhs, n * 4) != hipSuccess || hipMalloc(&N->w_solid, n * 4) != hipSuccess ||
        hipMalloc(&N->w_liquid, n * 4) != hipSuccess || hipMalloc(&N->w_pred, n * 4) != hipSuccess) {
        set_error("warm-start workspace: out of device memory");
        return VFLIP_ERR_NOMEM;
    }
    N->wx = cx;
    N->wy = cy;
    N->wz = cz;
    return 0;
}

}  // namespace

namespace vflip {
// The network's warm-start input buffers for a network grid of cx x cy x cz
// (restricted residual, solid and liquid fractions), allocated on first use.
int unet_warm_inputs(vflip_unet* N, int cx, int cy, int cz, float** crhs, float** csolid, float** cliquid) {
    VF_ARG(N && crhs && csolid && cliquid, "warm inputs");
    VF_ARG(dims_ok(cx, cy, cz), "network grid dims must be positive multiples of 8");
    if (int rc = ensure_warm_grid(N, cx, cy, cz)) return rc;
    if (int rc = ensure_grid(N, cx, cy, cz)) return rc;
    *crhs = N->w_rhs;
    *csolid = N->w_solid;
    *cliquid = N->w_liquid;
    return 0;
}

// The network on its filled input buffers (automatic normalisation), its prediction
// prolonged by `factor` and added into the FP64 guess g of the ni x nj x nk solve
// (cells with diag 0 unchanged). Asynchronous.
int unet_warm_predict_add(vflip_unet* N, int ni, int nj, int nk, int factor, const float* diag, double* g) {
    VF_ARG(N && g && factor >= 1, "warm predict args");
    const int cx = ni / factor, cy = nj / factor, cz = nk / factor;
    VF_ARG(cx * factor == ni && cy * factor == nj && cz * factor == nk && N->wx == cx && N->wy == cy && N->wz == cz,
           "warm inputs not prepared for this grid");
    const size_t Vc = (size_t)cx * cy * cz;
    const float* sd = prepare_scale(N, N->w_rhs, Vc, 0.f);
    k_pack_input<<<grid256(Vc), 256, 0, N->stream>>>(N->w_rhs, N->w_solid, N->w_liquid, Vc, 0.f, sd, N->act[kIn]);
    forward_layers(N);
    if (ni % 4 == 0 && ((uintptr_t)g & 31) == 0 && ((uintptr_t)diag & 15) == 0)
        k_prolong_add64_x4<<<dim3((ni / 4 + 255) / 256, nj, nk), 256, 0, N->stream>>>(N->pred, cx, cy, cz, factor, sd,
                                                                                      N->gain, diag, g);
    else
        k_prolong_add64<<<dim3((ni + 255) / 256, nj, nk), 256, 0, N->stream>>>(N->pred, cx, cy, cz, factor, sd,
                                                                              N->gain, diag, g);
    VF_CHECK_LAUNCH();
    return 0;
}
}  // namespace vflip

extern "C" {

int vflip_unet_create(unsigned seed, vflip_unet** out) {
    VF_ARG(out, "out");
    auto* N = new vflip_unet();
    size_t nw = 0, nb = 0, nwb = 0;
    for (int l = 0; l < kLayers; ++l) {
        N->woff[l] = nw;
        N->boff[l] = nb;
        N->wboff[l] = nwb;
        nw += (size_t)kNet[l].cout * kw_of(kNet[l]);
        nb += kNet[l].cout;
        nwb += (size_t)cin_of(kNet[l]) * kwb_of(kNet[l]);
    }
    N->nw = nw;
    N->nb = nb;
    N->nwb = nwb;
    const size_t np = nw + nb;
    if (hipMalloc(&N->params, np * 4) != hipSuccess || hipMalloc(&N->grads, np * 4) != hipSuccess ||
        hipMalloc(&N->adam_m, np * 4) != hipSuccess || hipMalloc(&N->adam_v, np * 4) != hipSuccess ||
        hipMalloc(&N->w16, nw * 2) != hipSuccess || hipMalloc(&N->wb16, nwb * 2) != hipSuccess ||
        hipMalloc(&N->scale, sizeof(float)) != hipSuccess || hipMalloc(&N->eacc, 2 * sizeof(double)) != hipSuccess) {
        delete N;
        set_error("U-Net parameters: out of device memory");
        return VFLIP_ERR_NOMEM;
    }
    // He initialisation over the real fan-in; padded entries stay zero
    std::vector<float> h(np, 0.f);
    Rng rng{0x5EEDull * 1000003ull + seed};
    for (int l = 0; l < kLayers; ++l) {
        const LayerDef& d = kNet[l];
        const int cin = cin_of(d), kw = kw_of(d);
        const int cout_real = l == kLayers - 1 ? 1 : d.cout;
        const double sd = std::sqrt(2.0 / (27.0 * d.cin_real)) * (l == kLayers - 1 ? 0.5 : 1.0);
        for (int co = 0; co < cout_real; ++co)
            for (int tap = 0; tap < 27; ++tap)
                for (int ci = 0; ci < cin; ++ci) {
                    if (l == 0 && ci >= d.cin_real) continue;
                    h[N->woff[l] + (size_t)co * kw + (size_t)tap * cin + ci] = (float)(sd * rng.normal());
                }
    }
    VF_HIP(hipMemcpy(N->params, h.data(), np * 4, hipMemcpyHostToDevice));
    VF_HIP(hipMemset(N->adam_m, 0, np * 4));
    VF_HIP(hipMemset(N->adam_v, 0, np * 4));
    VF_HIP(hipMemset(N->grads, 0, np * 4));
    k_to_bf16<<<grid256(nw), 256>>>(N->params, nw, N->w16);
    refresh_bwd_weights(N);
    VF_CHECK_LAUNCH();
    VF_HIP(hipDeviceSynchronize());
    *out = N;
    return 0;
}

int vflip_unet_destroy(vflip_unet* N) {
    if (!N) return 0;
    free_grid(N);
    for (void* p : {(void*)N->params, (void*)N->grads, (void*)N->adam_m, (void*)N->adam_v, (void*)N->w16,
                    (void*)N->wb16, (void*)N->scale, (void*)N->eacc, (void*)N->w_rhs, (void*)N->w_solid, (void*)N->w_liquid,
                    (void*)N->w_pred})
        if (p) (void)hipFree(p);
    delete N;
    return 0;
}

int vflip_unet_set_stream(vflip_unet* N, void* hip_stream) {
    VF_ARG(N, "handle");
    N->stream = (hipStream_t)hip_stream;
    return 0;
}

size_t vflip_unet_num_params(const vflip_unet* N) { return N ? N->nw + N->nb : 0; }

int vflip_unet_set_gain(vflip_unet* N, float gain) {
    VF_ARG(N && gain > 0.f, "gain");
    N->gain = gain;
    return 0;
}

float vflip_unet_gain(const vflip_unet* N) { return N ? N->gain : 0.f; }

int vflip_unet_get_params(const vflip_unet* N, float* dst) {
    VF_ARG(N && dst, "args");
    VF_HIP(hipMemcpyAsync(dst, N->params, (N->nw + N->nb) * 4, hipMemcpyDeviceToDevice, N->stream));
    VF_HIP(hipStreamSynchronize(N->stream));
    return 0;
}

int vflip_unet_set_params(vflip_unet* N, const float* src) {
    VF_ARG(N && src, "args");
    VF_HIP(hipMemcpyAsync(N->params, src, (N->nw + N->nb) * 4, hipMemcpyDeviceToDevice, N->stream));
    k_to_bf16<<<grid256(N->nw), 256, 0, N->stream>>>(N->params, N->nw, N->w16);
    refresh_bwd_weights(N);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(N->stream));
    return 0;
}

int vflip_unet_get_grads(const vflip_unet* N, float* dst) {
    VF_ARG(N && dst, "args");
    VF_HIP(hipMemcpyAsync(dst, N->grads, (N->nw + N->nb) * 4, hipMemcpyDeviceToDevice, N->stream));
    VF_HIP(hipStreamSynchronize(N->stream));
    return 0;
}

int vflip_unet_get_activation(const vflip_unet* N, int tensor, float* dst) {
    VF_ARG(N && dst, "args");
    VF_ARG(tensor >= 0 && tensor < kTensors && tensor != kLayers - 1, "tensor (0..13 layer outputs, 15..17 pooled, 18 input)");
    const bf16* a = N->act[tensor];
    VF_ARG(a != nullptr, "no forward has run on a grid");
    const size_t n = level_vox(N, tensor_level(tensor)) * tensor_channels(tensor);
    k_bf16_to_f32<<<grid256(n), 256, 0, N->stream>>>(a, n, dst);
    VF_CHECK_LAUNCH();
    VF_HIP(hipStreamSynchronize(N->stream));
    return 0;
}

int vflip_unet_zero_grad(vflip_unet* N) {
    VF_ARG(N, "handle");
    VF_HIP(hipMemsetAsync(N->grads, 0, (N->nw + N->nb) * 4, N->stream));
    return 0;
}

int vflip_unet_forward(vflip_unet* N, int ni, int nj, int nk, const float* rhs, const float* solid,
                       const float* liquid, float in_scale, float* pred) {
    VF_ARG(N && rhs && pred, "forward args");
    VF_ARG(dims_ok(ni, nj, nk), "grid dims must be positive multiples of 8");
    if (int rc = ensure_grid(N, ni, nj, nk)) return rc;
    const size_t V = (size_t)ni * nj * nk;
    const float* sd = prepare_scale(N, rhs, V, in_scale);
    k_pack_input<<<grid256(V), 256, 0, N->stream>>>(rhs, solid, liquid, V, in_scale, sd, N->act[kIn]);
    forward_layers(N);
    k_scale_copy<<<grid256(V), 256, 0, N->stream>>>(N->pred, V, sd, N->gain, pred);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_unet_backward(vflip_unet* N, const float* grad_pred) {
    VF_ARG(N && grad_pred && N->nx > 0, "backward needs a forward first");
    const size_t V = level_vox(N, 0);
    VF_HIP(hipMemcpyAsync(N->gpred, grad_pred, V * 4, hipMemcpyDeviceToDevice, N->stream));
    backward_layers(N);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_unet_adam(vflip_unet* N, float lr, int step, float grad_scale) {
    VF_ARG(N && step >= 1, "adam args");
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)step)), c2 = 1.f / (1.f - std::pow(b2, (float)step));
    k_adam<<<grid256(N->nw), 256, 0, N->stream>>>(N->params, N->grads, N->adam_m, N->adam_v, N->nw, lr, b1, b2, eps,
                                                   c1, c2, grad_scale, N->w16);
    k_adam<<<grid256(N->nb), 256, 0, N->stream>>>(N->params + N->nw, N->grads + N->nw, N->adam_m + N->nw,
                                                   N->adam_v + N->nw, N->nb, lr, b1, b2, eps, c1, c2, grad_scale,
                                                   nullptr);
    refresh_bwd_weights(N);
    VF_CHECK_LAUNCH();
    return 0;
}

int vflip_unet_train_step(vflip_unet* N, int ni, int nj, int nk, const float* rhs, const float* solid,
                          const float* liquid, float in_scale, const float* target, double* loss) {
    VF_ARG(N && rhs && target, "train args");
    VF_ARG(dims_ok(ni, nj, nk), "grid dims must be positive multiples of 8");
    if (int rc = ensure_grid(N, ni, nj, nk)) return rc;
    const size_t V = (size_t)ni * nj * nk;
    const float* sd = prepare_scale(N, rhs, V, in_scale);
    k_pack_input<<<grid256(V), 256, 0, N->stream>>>(rhs, solid, liquid, V, in_scale, sd, N->act[kIn]);
    forward_layers(N);
    VF_HIP(hipMemsetAsync(N->loss, 0, sizeof(double), N->stream));
    k_mse<<<grid256(V), 256, 0, N->stream>>>(N->pred, target, V, sd, N->gain, N->gpred, N->loss);
    backward_layers(N);
    VF_CHECK_LAUNCH();
    if (loss) {
        VF_HIP(hipMemcpyAsync(loss, N->loss, sizeof(double), hipMemcpyDeviceToHost, N->stream));
        VF_HIP(hipStreamSynchronize(N->stream));
    }
    return 0;
}

// Energy-norm training step: forward, loss = |pred - p/(s gain)|_A^2 / |p/(s gain)|_A^2
// with the sample's operator A (automatic normalisation), backward.
int vflip_unet_train_step_energy(vflip_unet* N, int ni, int nj, int nk, const float* rhs, const float* solid,
                                 const float* liquid, const float* target, const float* diag, const float* ax,
                                 const float* ay, const float* az, double* loss) {
    VF_ARG(N && rhs && target && diag && ax && ay && az, "train args");
    VF_ARG(dims_ok(ni, nj, nk), "grid dims must be positive multiples of 8");
    if (int rc = ensure_grid(N, ni, nj, nk)) return rc;
    const size_t V = (size_t)ni * nj * nk;
    const float* sd = prepare_scale(N, rhs, V, 0.f);
    k_pack_input<<<grid256(V), 256, 0, N->stream>>>(rhs, solid, liquid, V, 0.f, sd, N->act[kIn]);
    forward_layers(N);
    // scratch: e and x in gx, A e after them (gx holds at least 16 V floats)
    float* e = N->gx;
    float* xn = N->gx + V;
    float* ae = N->gx + 2 * V;
    const Stencil A{diag, ax, ay, az, ni, nj, nk};
    k_energy_prep<<<grid256(V), 256, 0, N->stream>>>(N->pred, target, V, sd, N->gain, e, xn, N->eacc);
    k_stencil_quad<<<grid256(V), 256, 0, N->stream>>>(A, e, ae, N->eacc, 0);
    k_stencil_quad<<<grid256(V), 256, 0, N->stream>>>(A, xn, nullptr, N->eacc, 1);
    k_energy_grad<<<grid256(V), 256, 0, N->stream>>>(ae, V, N->eacc, N->gpred, N->loss);
    backward_layers(N);
    VF_CHECK_LAUNCH();
    if (loss) {
        VF_HIP(hipMemcpyAsync(loss, N->loss, sizeof(double), hipMemcpyDeviceToHost, N->stream));
        VF_HIP(hipStreamSynchronize(N->stream));
    }
    return 0;
}

// Warm start for a solve on an ni x nj x nk grid: inputs restricted by
// `factor` to the network grid, network in automatic normalisation, pressure
// prolonged back (trilinear) into x0. diag: the solve's level-0 diagonal
// (0 marks solid and Dirichlet cells; x0 is 0 there). Asynchronous.
int vflip_unet_warm_start(vflip_unet* N, int ni, int nj, int nk, int factor, const float* rhs, const float* diag,
                          float* x0) {
    VF_ARG(N && rhs && x0 && factor >= 1, "warm_start args");
    VF_ARG(ni % factor == 0 && nj % factor == 0 && nk % factor == 0, "grid dims divisible by factor");
    const int cx = ni / factor, cy = nj / factor, cz = nk / factor;
    VF_ARG(dims_ok(cx, cy, cz), "network grid dims must be positive multiples of 8");
    if (int rc = ensure_warm_grid(N, cx, cy, cz)) return rc;
    if (int rc = ensure_grid(N, cx, cy, cz)) return rc;
    const size_t Vc = (size_t)cx * cy * cz;
    auto a16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
    const int vec4 = factor == 4 && ni % 4 == 0 && a16(rhs) && (!diag || a16(diag));
    k_restrict_inputs<<<dim3((cx + 255) / 256, cy, cz), 256, 0, N->stream>>>(rhs, diag, ni, nj, nk, factor, vec4,
                                                                             N->w_rhs, N->w_solid, N->w_liquid);
    const float* sd = prepare_scale(N, N->w_rhs, Vc, 0.f);
    k_pack_input<<<grid256(Vc), 256, 0, N->stream>>>(N->w_rhs, N->w_solid, N->w_liquid, Vc, 0.f, sd, N->act[kIn]);
    forward_layers(N);
    const size_t V = (size_t)ni * nj * nk;
    k_prolong_cells<<<dim3((ni + 255) / 256, nj, nk), 256, 0, N->stream>>>(N->pred, cx, cy, cz, factor, sd, N->gain,
                                                                           diag, x0);
    VF_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
