// tcnn_hashgrid.hip — the jax-tcnn hash-grid encoding (inference-time encoder of
// the reference: models/encoders.py:259-305 TCNNHashGridEncoder over
// deps/jax-tcnn/lib/impl/hashgrid.cu, which launches tiny-cuda-nn v1.6's
// kernel_grid / kernel_grid_backward / kernel_grid_backward_input).
//
// tiny-cuda-nn is not vendored with the reference; its published grid encoding is
// restated here, with the layouts jax-tcnn binds:
//   coords [3, N] (feature-major), table [offsets[L], F] with unpadded level sizes
//   (res^3 for dense levels, T otherwise), encodings [L*F, N], dy_dcoords holding
//   for feature k and point i the 3-vector d enc[k, i] / d coords[:, i] at
//   ((k * N) + i) * 3.
// Per level l: scale = exp2(l * log2(per_level_scale)) * N_min - 1 (f32; exp2
// rounded once here, from double), resolution = ceil(scale) + 1; per coordinate
// pos = fma(scale, x, 0.5), grid = floor(pos), frac = pos - grid; corner c (bit d
// of c = dimension d, x lowest) has weight prod_d (bit ? frac_d : 1 - frac_d) (in
// dimension order) and index: dense (x + y res + z res^2 while the stride stays
// within the level's size) or the coherent prime hash x ^ y 2654435761 ^ z 805459861,
// either taken modulo the level's size; enc = sum_c fma(weight_c, row_c, enc) in
// corner order.  The gradient wrt coords, per dimension d: the four edges along d,
// weight = scale * prod of the other two dimensions' factors, times (right - left).
//
// Differences from jax-tcnn: the level offsets are read on the device (jax-tcnn
// copies them to the host and reads them without waiting for the copy), errors are
// reported through the library's error state instead of thrown, and the table
// gradient is zeroed by a kernel that reads its size on the device (no host copy).
#include "ngp_common.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMaxLevels = 32;

struct TcnnLevels {
    float scale[kMaxLevels];
    uint32_t res[kMaxLevels];
};

__device__ __forceinline__ void tcnn_levels(uint32_t L, uint32_t N_min, float log2_pls, TcnnLevels &lv) {
    for (uint32_t l = 0; l < L; ++l) {
        const float t = (float)l * log2_pls;
        const float e = (float)exp2((double)t);
        lv.scale[l] = e * (float)N_min - 1.0f;
        lv.res[l] = (uint32_t)ceilf(lv.scale[l]) + 1u;
    }
}

__device__ __forceinline__ uint32_t tcnn_index(uint32_t size, uint32_t res, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t stride = 1, index = 0;
    const uint32_t p[3] = {x, y, z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (stride > size) break;
        index += p[d] * stride;
        stride *= res;
    }
    if (size < stride) index = x ^ (y * 2654435761u) ^ (z * 805459861u);
    return index % size;
}

struct TcnnPoint {
    float frac[3];
    uint32_t grid[3];
};

__device__ __forceinline__ TcnnPoint tcnn_point(const float *__restrict__ coords, uint32_t n, uint32_t i, float scale) {
    TcnnPoint p;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float pos = fmaf(scale, coords[(size_t)d * n + i], 0.5f);
        const float g = floorf(pos);
        p.grid[d] = (uint32_t)(int)g;
        p.frac[d] = pos - g;
    }
    return p;
}

template <int F>
__global__ void __launch_bounds__(kBlock)
    tcnn_forward_kernel(uint32_t n, uint32_t L, uint32_t N_min, float log2_pls, const uint32_t *__restrict__ offsets,
                        const float *__restrict__ coords, const float *__restrict__ table, float *__restrict__ enc,
                        float *__restrict__ dy_dx) {
    __shared__ TcnnLevels lv;
    __shared__ uint32_t offs[kMaxLevels + 1];
    if (threadIdx.x == 0) tcnn_levels(L, N_min, log2_pls, lv);
    if (threadIdx.x <= L) offs[threadIdx.x] = offsets[threadIdx.x];
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t l = blockIdx.y;
    if (i >= n) return;
    if (offs[l + 1] <= offs[l]) {  // an empty (or inverted) level encodes to zero
#pragma unroll
        for (int f = 0; f < F; ++f) enc[(size_t)(l * F + f) * n + i] = 0.0f;
        if (dy_dx != nullptr)
            for (int k = 0; k < 3 * F; ++k) dy_dx[((size_t)(l * F + k / 3) * n + i) * 3 + k % 3] = 0.0f;
        return;
    }
    const uint32_t size = offs[l + 1] - offs[l];
    const float *grid = table + (size_t)offs[l] * F;
    const float scale = lv.scale[l];
    const uint32_t res = lv.res[l];
    const TcnnPoint p = tcnn_point(coords, n, i, scale);
    auto row = [&](uint32_t x, uint32_t y, uint32_t z) { return grid + (size_t)tcnn_index(size, res, x, y, z) * F; };
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float w = 1.0f;
        uint32_t q[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const bool hi = (c >> d) & 1;
            w = w * (hi ? p.frac[d] : 1.0f - p.frac[d]);
            q[d] = p.grid[d] + (hi ? 1u : 0u);
        }
        const float *r = row(q[0], q[1], q[2]);
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = fmaf(w, r[f], acc[f]);
    }
#pragma unroll
    for (int f = 0; f < F; ++f) enc[(size_t)(l * F + f) * n + i] = acc[f];
    if (dy_dx == nullptr) return;
#pragma unroll
    for (int gd = 0; gd < 3; ++gd) {
        float grad[F];
#pragma unroll
        for (int f = 0; f < F; ++f) grad[f] = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float w = scale;
            uint32_t q[3];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int d = k >= gd ? k + 1 : k;
                const bool hi = (c >> k) & 1;
                w = w * (hi ? p.frac[d] : 1.0f - p.frac[d]);
                q[d] = p.grid[d] + (hi ? 1u : 0u);
            }
            q[gd] = p.grid[gd];
            const float *left = row(q[0], q[1], q[2]);
            q[gd] = p.grid[gd] + 1u;
            const float *right = row(q[0], q[1], q[2]);
#pragma unroll
            for (int f = 0; f < F; ++f) grad[f] = grad[f] + w * (right[f] - left[f]);
        }
#pragma unroll
        for (int f = 0; f < F; ++f) dy_dx[((size_t)(l * F + f) * n + i) * 3 + gd] = grad[f];
    }
}

template <int F>
__global__ void __launch_bounds__(kBlock)
    tcnn_backward_kernel(uint32_t n, uint32_t L, uint32_t N_min, float log2_pls, const uint32_t *__restrict__ offsets,
                         const float *__restrict__ coords, const float *__restrict__ dL_dy,
                         float *__restrict__ dtable) {
    __shared__ TcnnLevels lv;
    __shared__ uint32_t offs[kMaxLevels + 1];
    if (threadIdx.x == 0) tcnn_levels(L, N_min, log2_pls, lv);
    if (threadIdx.x <= L) offs[threadIdx.x] = offsets[threadIdx.x];
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t l = blockIdx.y;
    if (i >= n || offs[l + 1] <= offs[l]) return;
    const uint32_t size = offs[l + 1] - offs[l];
    float *grid = dtable + (size_t)offs[l] * F;
    const float scale = lv.scale[l];
    const uint32_t res = lv.res[l];
    const TcnnPoint p = tcnn_point(coords, n, i, scale);
    float g[F];
#pragma unroll
    for (int f = 0; f < F; ++f) g[f] = dL_dy[(size_t)(l * F + f) * n + i];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float w = 1.0f;
        uint32_t q[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const bool hi = (c >> d) & 1;
            w = w * (hi ? p.frac[d] : 1.0f - p.frac[d]);
            q[d] = p.grid[d] + (hi ? 1u : 0u);
        }
        float *r = grid + (size_t)tcnn_index(size, res, q[0], q[1], q[2]) * F;
#pragma unroll
        for (int f = 0; f < F; ++f) atomicAdd(r + f, w * g[f]);
    }
}

// dL/dcoords[d, i] = sum_k dL_dy[k, i] * dy_dx[k, i, d], k in feature order.
__global__ void __launch_bounds__(kBlock)
    tcnn_backward_input_kernel(uint32_t n, uint32_t n_features, const float *__restrict__ dL_dy,
                               const float *__restrict__ dy_dx, float *__restrict__ dL_dx) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float r[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < n_features; ++k) {
        const float gk = dL_dy[(size_t)k * n + i];
        const float *v = dy_dx + ((size_t)k * n + i) * 3;
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] = fmaf(gk, v[d], r[d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) dL_dx[(size_t)d * n + i] = r[d];
}

// Zeroes the table gradient, offsets[L] * F floats: the size is read on the device
// (a fixed grid strides over it), so the call needs no host round trip.
__global__ void zero_table_kernel(const uint32_t *__restrict__ offsets_end, uint32_t F, float *__restrict__ a) {
    const size_t n = (size_t)offsets_end[0] * F;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) a[j] = 0.0f;
}

bool validate(const TcnnHashGridDescriptor &d, const char *target) {
    if (d.L == 0 || d.L > kMaxLevels || !(d.F == 2 || d.F == 4) || d.N_min == 0 || !(d.per_level_scale > 0.0f)) {
        set_error(kBadDescriptor, "%s: need 1 <= L <= %u, F in {2, 4}, N_min > 0 and per_level_scale > 0 (L=%u F=%u)",
                  target, kMaxLevels, d.L, d.F);
        return false;
    }
    return true;
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void tcnn_hashgrid_encode(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "tcnn_hashgrid_encode";
    TcnnHashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!validate(d, target)) return;
    if (d.n_coords == 0) return;
    if (!check_buffers(buffers, 5, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    const dim3 grid(div_up(d.n_coords, kBlock), d.L);
    const float l2 = log2f(d.per_level_scale);
    if (d.F == 2) {
        hipLaunchKernelGGL(tcnn_forward_kernel<2>, grid, dim3(kBlock), 0, stream, d.n_coords, d.L, d.N_min, l2,
                           (const uint32_t *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                           (float *)buffers[3], (float *)buffers[4]);
    } else {
        hipLaunchKernelGGL(tcnn_forward_kernel<4>, grid, dim3(kBlock), 0, stream, d.n_coords, d.L, d.N_min, l2,
                           (const uint32_t *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                           (float *)buffers[3], (float *)buffers[4]);
    }
    check_launch(target);
}

extern "C" void tcnn_hashgrid_encode_backward(ngp_stream_t stream_, void **buffers, const char *opaque,
                                              size_t opaque_len) {
    const char *target = "tcnn_hashgrid_encode_backward";
    TcnnHashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!validate(d, target)) return;
    if (!check_buffers(buffers, 6, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    hipLaunchKernelGGL(zero_table_kernel, dim3(2048), dim3(kBlock), 0, stream, (const uint32_t *)buffers[0] + d.L, d.F,
                       (float *)buffers[4]);
    if (d.n_coords > 0) {
        const dim3 grid(div_up(d.n_coords, kBlock), d.L);
        const float l2 = log2f(d.per_level_scale);
        if (d.F == 2) {
            hipLaunchKernelGGL(tcnn_backward_kernel<2>, grid, dim3(kBlock), 0, stream, d.n_coords, d.L, d.N_min, l2,
                               (const uint32_t *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                               (float *)buffers[4]);
        } else {
            hipLaunchKernelGGL(tcnn_backward_kernel<4>, grid, dim3(kBlock), 0, stream, d.n_coords, d.L, d.N_min, l2,
                               (const uint32_t *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                               (float *)buffers[4]);
        }
        hipLaunchKernelGGL(tcnn_backward_input_kernel, dim3(div_up(d.n_coords, kBlock)), dim3(kBlock), 0, stream,
                           d.n_coords, d.L * d.F, (const float *)buffers[2], (const float *)buffers[3],
                           (float *)buffers[5]);
    }
    check_launch(target);
}
