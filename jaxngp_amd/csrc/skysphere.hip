// skysphere.hip — the sky-sphere background model of unbounded scenes
// (models/nerfs.py:133-213, shapes from make_skysphere_background_model_ngp,
// nerfs.py:383-419).
//
//   per ray:  a = d.d, b = 2 o.d, c = o.o - r^2 (nerfs.py:172-174), t = the larger
//             root of a t^2 + b t + c = 0 with NaN propagating like jnp.maximum
//             (nerfs.py:181-184), pos = o + t d;
//   batch:    D = ||pos||_F + 1e-15 over the whole [n, 3] array (nerfs.py:191-196:
//             jnp.linalg.norm without an axis), so one ray that misses the sphere
//             (NaN t) makes every colour NaN, as in the reference;
//   per ray:  x = [SH2(finite(t) ? pos / D : 0), SH4(d), appearance] (20 + E),
//             colour = sigmoid(relu(relu(x W0) W1) W2) where t is finite, else 0.
//
// Kernels: two fixed-order passes for D (f64 partial sums, one block each over a
// fixed grid, then one block adds them), the forward (a thread per ray; the
// weights are read at wave-uniform addresses, so they arrive by scalar loads and
// enter the FMAs as scalar operands) and the backward (a thread per ray for the
// per-ray chain; the weight gradient sum_rays act (x) delta of each layer is staged
// through LDS, 256 rays at a time, and accumulated by register tiles of 4-5 rows x
// 1 column per thread; one workspace row per workgroup, rows added in a fixed
// order).  The MLP is small (2,272 MACs per ray at most) and off the sampled-point
// path (one evaluation per ray), so VALU FMAs suffice; MFMA tiles would be mostly
// padding at K = 20.
#include "ngp_common.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;
constexpr int kNormBlocks = 256;     // fixed grid of the norm's partial sums
constexpr int kMaxBwdBlocks = 512;   // backward workgroups (two per CU)
constexpr int kPos = 4, kDir = 16;   // SH degree 2 of the hit point, degree 4 of d
constexpr int kIn0 = kPos + kDir;    // 20 inputs before the appearance embedding
constexpr int kHid = 32, kOut = 3;
constexpr int kMaxE = NGP_SKY_MAX_EXTRA_DIMS;
constexpr int kMaxIn = kIn0 + kMaxE;
constexpr size_t kDenomOffset = sizeof(double) * kNormBlocks;  // 2048
constexpr size_t kRowsOffset = kDenomOffset + 256;
static_assert(kBlock == kNormBlocks, "the final norm pass loads one partial per thread");
static_assert(kBlock == 8 * kHid, "backward register tiles: 8 row groups x 32 columns");

__host__ __device__ inline uint32_t n_params(uint32_t E) { return (kIn0 + E) * kHid + kHid * kHid + kHid * kOut; }

uint32_t bwd_blocks(uint32_t n) {
    return std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, kBlock), kMaxBwdBlocks));
}

struct SkyArgs {
    uint32_t n, E, per_ray;
    float r2;  // radius^2 rounded to f32 (the reference's self.r ** 2 meets f32 arrays)
};

__device__ __forceinline__ float nan_max(float a, float b) {
    // jnp.maximum: NaN if either operand is NaN
    return (a != a || b != b) ? __builtin_nanf("") : (a > b ? a : b);
}

// relu as jnp.maximum(x, 0): NaN stays NaN
__device__ __forceinline__ float relu(float x) { return x <= 0.0f ? 0.0f : x; }

// Far intersection of o + t d with the sphere (nerfs.py:172-190); every f32
// operation rounded on its own, sums in the order jnp's 3-element reductions take.
__device__ __forceinline__ float far_hit(const float *__restrict__ o, const float *__restrict__ d, size_t i,
                                         float r2, float p[3], float dv[3]) {
    const float ox = o[3 * i], oy = o[3 * i + 1], oz = o[3 * i + 2];
    const float dx = d[3 * i], dy = d[3 * i + 1], dz = d[3 * i + 2];
    const float a = (dx * dx + dy * dy) + dz * dz;
    const float b = 2.0f * ((ox * dx + oy * dy) + oz * dz);
    const float c = ((ox * ox + oy * oy) + oz * oz) - r2;
    const float s = sqrtf(b * b - (4.0f * a) * c);
    const float t = nan_max((-b + s) / (2.0f * a), (-b - s) / (2.0f * a));
    p[0] = ox + t * dx;
    p[1] = oy + t * dy;
    p[2] = oz + t * dz;
    dv[0] = dx;
    dv[1] = dy;
    dv[2] = dz;
    return t;
}

// SH degree 4 of (x, y, z) into e[0..15] (encoders.py:376-405); degree 2 is e[0..3].
__device__ __forceinline__ void sh16(float x, float y, float z, float *e) {
    const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
    e[0] = 0.28209479177387814f;
    e[1] = -0.48860251190291987f * y;
    e[2] = 0.48860251190291987f * z;
    e[3] = -0.48860251190291987f * x;
    e[4] = 1.0925484305920792f * xy;
    e[5] = -1.0925484305920792f * yz;
    e[6] = 0.94617469575755997f * z2 - 0.31539156525251999f;
    e[7] = -1.0925484305920792f * xz;
    e[8] = 0.54627421529603959f * x2 - 0.54627421529603959f * y2;
    e[9] = 0.59004358992664352f * y * (-3.0f * x2 + y2);
    e[10] = 2.8906114426405538f * xy * z;
    e[11] = 0.45704579946446572f * y * (1.0f - 5.0f * z2);
    e[12] = 0.3731763325901154f * z * (5.0f * z2 - 3.0f);
    e[13] = 0.45704579946446572f * x * (1.0f - 5.0f * z2);
    e[14] = 1.4453057213202769f * z * (x2 - y2);
    e[15] = 0.59004358992664352f * x * (-x2 + 3.0f * y2);
}

// The MLP input of ray i: [SH2(pos / D or 0), SH4(d), appearance].  Returns finite(t).
__device__ __forceinline__ bool ray_input(const SkyArgs &a, const float *__restrict__ o, const float *__restrict__ d,
                                          const float *__restrict__ app, float denom, uint32_t i, float *x) {
    float p[3], dv[3];
    const float t = far_hit(o, d, i, a.r2, p, dv);
    const bool fin = __builtin_isfinite(t);
    float e[16];
    sh16(fin ? p[0] / denom : 0.0f, fin ? p[1] / denom : 0.0f, fin ? p[2] / denom : 0.0f, e);
#pragma unroll
    for (int k = 0; k < kPos; ++k) x[k] = e[k];
    sh16(dv[0], dv[1], dv[2], e);
#pragma unroll
    for (int k = 0; k < kDir; ++k) x[kPos + k] = e[k];
    const float *ar = app + (a.per_ray ? (size_t)i * a.E : 0);
#pragma unroll
    for (int k = 0; k < kMaxE; ++k) x[kIn0 + k] = k < (int)a.E ? ar[k] : 0.0f;
    return fin;
}

// h1 = relu(x W0), h2 = relu(h1 W1), o = h2 W2; dot products in ascending input
// order.  The weights are read at wave-uniform addresses of the read-only params
// array, so they come through scalar loads and feed the FMAs as scalar operands
// (no vector registers or LDS traffic for them).  relu'(z) = (z > 0) is read off
// h = relu(z) as (h > 0), which agrees for every z including NaN.
__device__ __forceinline__ void mlp_forward(const float *__restrict__ w, uint32_t K, const float *x, float *h1,
                                            float *h2, float *o) {
#pragma unroll
    for (int j = 0; j < kHid; ++j) h1[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < kMaxIn; ++k) {
        if (k < (int)K) {
#pragma unroll
            for (int j = 0; j < kHid; ++j) h1[j] = h1[j] + x[k] * w[k * kHid + j];
        }
    }
#pragma unroll
    for (int j = 0; j < kHid; ++j) h1[j] = relu(h1[j]);
    const float *w1 = w + K * kHid;
#pragma unroll
    for (int j = 0; j < kHid; ++j) h2[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < kHid; ++k) {
#pragma unroll
        for (int j = 0; j < kHid; ++j) h2[j] = h2[j] + h1[k] * w1[k * kHid + j];
    }
#pragma unroll
    for (int j = 0; j < kHid; ++j) h2[j] = relu(h2[j]);
    const float *w2 = w1 + kHid * kHid;
    o[0] = o[1] = o[2] = 0.0f;
#pragma unroll
    for (int k = 0; k < kHid; ++k) {
#pragma unroll
        for (int j = 0; j < kOut; ++j) o[j] = o[j] + h2[k] * w2[k * kOut + j];
    }
}

__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + det_expf(-x)); }

// sum over all rays of |pos|^2 in f64, one partial per workgroup of a fixed grid
__global__ void __launch_bounds__(kBlock)
    sky_norm_partial_kernel(SkyArgs a, const float *__restrict__ o, const float *__restrict__ d,
                            double *__restrict__ partial) {
    __shared__ double wsum[kBlock / kWave];
    double s = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += kNormBlocks * kBlock) {
        float p[3], dv[3];
        far_hit(o, d, i, a.r2, p, dv);
        s += (double)p[0] * p[0] + (double)p[1] * p[1] + (double)p[2] * p[2];
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += wsum[w];
        partial[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kBlock)
    sky_norm_final_kernel(const double *__restrict__ partial, float *__restrict__ denom) {
    __shared__ double wsum[kBlock / kWave];
    double s = partial[threadIdx.x];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += wsum[w];
        *denom = (float)sqrt(t) + 1e-15f;
    }
}

__global__ void __launch_bounds__(kBlock)
    sky_forward_kernel(SkyArgs a, const float *__restrict__ o, const float *__restrict__ d,
                       const float *__restrict__ app, const float *__restrict__ params,
                       const float *__restrict__ denom_p, float *__restrict__ colors) {
    const uint32_t K = kIn0 + a.E;
    const float denom = *denom_p;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < a.n; i += gridDim.x * kBlock) {
        float x[kMaxIn], h1[kHid], h2[kHid], out[kOut];
        const bool fin = ray_input(a, o, d, app, denom, i, x);
        mlp_forward(params, K, x, h1, h2, out);
#pragma unroll
        for (int j = 0; j < kOut; ++j) colors[3 * (size_t)i + j] = fin ? sigmoid(out[j]) : 0.0f;
    }
}

// Backward.  Per tile of 256 rays (one per thread) three LDS stages, one per
// layer, hold each ray's (activation, delta) pair; thread (q, j) = (tid / 32,
// tid % 32) accumulates rows 4q..4q+3 of dW1 and rows 5q..5q+4 of dW0 in column
// j, threads below 96 one element (tid / 3, tid % 3) of dW2, over the tile's rays
// in ray order.
constexpr int kRowC = kHid + kOut;         // [h2 | do]   35 (odd)
constexpr int kRowB = 2 * kHid + 1;        // [h1 | dz2]  65 (odd)
constexpr int kRowAMax = kMaxIn + kHid + 1;  // [x | dz1] padded to odd
constexpr int kStage = kBlock * kRowAMax;

__global__ void __launch_bounds__(kBlock)
    sky_backward_kernel(SkyArgs a, const float *__restrict__ o, const float *__restrict__ d,
                        const float *__restrict__ app, const float *__restrict__ params,
                        const float *__restrict__ denom_p, const float *__restrict__ dcolors,
                        float *__restrict__ dapp, float *__restrict__ rows) {
    __shared__ float stage[kStage];
    const uint32_t K = kIn0 + a.E;
    const uint32_t P = n_params(a.E);
    const uint32_t rowA = (K + kHid) | 1u;
    const float *w1 = params + K * kHid;
    const float *w2 = w1 + kHid * kHid;
    const float denom = *denom_p;
    const int tid = (int)threadIdx.x;
    const int q = tid / kHid, jc = tid % kHid;
    float acc0[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, acc1[4] = {0.0f, 0.0f, 0.0f, 0.0f}, acc2 = 0.0f;
    for (uint32_t base = blockIdx.x * kBlock; base < a.n; base += gridDim.x * kBlock) {
        const uint32_t i = base + tid;
        const bool live = i < a.n;
        float x[kMaxIn], h1[kHid], h2[kHid], dout[kOut];
        if (live) {
            float out[kOut];
            const bool fin = ray_input(a, o, d, app, denom, i, x);
            mlp_forward(params, K, x, h1, h2, out);
#pragma unroll
            for (int j = 0; j < kOut; ++j) {
                // jnp.where passes the cotangent only where t is finite; then the
                // logistic's VJP g y (1 - y)
                const float g = fin ? dcolors[3 * (size_t)i + j] : 0.0f;
                const float y = sigmoid(out[j]);
                dout[j] = g * y * (1.0f - y);
            }
        } else {
#pragma unroll
            for (int k = 0; k < kMaxIn; ++k) x[k] = 0.0f;
#pragma unroll
            for (int j = 0; j < kHid; ++j) h1[j] = h2[j] = 0.0f;
#pragma unroll
            for (int j = 0; j < kOut; ++j) dout[j] = 0.0f;
        }
        // stage C: dW2 += h2 (x) do
        float *row = stage + tid * kRowC;
#pragma unroll
        for (int k = 0; k < kHid; ++k) row[k] = h2[k];
#pragma unroll
        for (int j = 0; j < kOut; ++j) row[kHid + j] = dout[j];
        __syncthreads();
        if (tid < kHid * kOut) {
            const int k2 = tid / kOut, j2 = tid % kOut;
            for (int r = 0; r < kBlock; ++r) acc2 = acc2 + stage[r * kRowC + k2] * stage[r * kRowC + kHid + j2];
        }
        __syncthreads();
        // dz2 = relu'(z2) (W2 do), written straight into stage B next to h1
        row = stage + tid * kRowB;
#pragma unroll
        for (int k = 0; k < kHid; ++k) {
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < kOut; ++j) s = s + w2[k * kOut + j] * dout[j];
            row[kHid + k] = h2[k] > 0.0f ? s : 0.0f;
            row[k] = h1[k];
        }
        __syncthreads();
        for (int r = 0; r < kBlock; ++r) {
            const float *sr = stage + r * kRowB;
            const float dj = sr[kHid + jc];
#pragma unroll
            for (int m = 0; m < 4; ++m) acc1[m] = acc1[m] + sr[4 * q + m] * dj;
        }
        // dz1 = relu'(z1) (W1 dz2), dz2 read back from this thread's own row
        float dz1[kHid];
#pragma unroll
        for (int k = 0; k < kHid; ++k) dz1[k] = 0.0f;
        {
            const float *mine = stage + tid * kRowB + kHid;
#pragma unroll
            for (int j = 0; j < kHid; ++j) {
                const float dzj = mine[j];
#pragma unroll
                for (int k = 0; k < kHid; ++k) dz1[k] = dz1[k] + w1[k * kHid + j] * dzj;
            }
        }
#pragma unroll
        for (int k = 0; k < kHid; ++k) dz1[k] = h1[k] > 0.0f ? dz1[k] : 0.0f;
        __syncthreads();
        // stage A: dW0 += x (x) dz1
        row = stage + tid * rowA;
#pragma unroll
        for (int k = 0; k < kMaxIn; ++k) {
            if (k < (int)K) row[k] = x[k];
        }
#pragma unroll
        for (int k = 0; k < kHid; ++k) row[K + k] = dz1[k];
        __syncthreads();
        for (int r = 0; r < kBlock; ++r) {
            const float *sr = stage + r * rowA;
            const float dj = sr[K + jc];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const int k = 5 * q + m;
                if (k < (int)K) acc0[m] = acc0[m] + sr[k] * dj;
            }
        }
        __syncthreads();
        // dL/dappearance = (W0 dz1) on the appearance rows
        if (live) {
            for (uint32_t e = 0; e < a.E; ++e) {
                const float *wr = params + (kIn0 + e) * kHid;
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < kHid; ++j) s = s + wr[j] * dz1[j];
                dapp[(size_t)i * a.E + e] = s;
            }
        }
    }
    float *out = rows + (size_t)blockIdx.x * P;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const int k = 5 * q + m;
        if (k < (int)K) out[k * kHid + jc] = acc0[m];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) out[K * kHid + (4 * q + m) * kHid + jc] = acc1[m];
    if (tid < kHid * kOut) out[K * kHid + kHid * kHid + tid] = acc2;
}

__global__ void __launch_bounds__(kBlock)
    sky_sum_rows_kernel(uint32_t n_rows, uint32_t P, const float *__restrict__ rows, float *__restrict__ out) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= P) return;
    double s = 0.0;
    for (uint32_t r = 0; r < n_rows; ++r) s += rows[(size_t)r * P + k];
    out[k] = (float)s;
}

bool valid(const SkySphereBgDescriptor &d, const char *target) {
    if (d.n_extra_dims > kMaxE || d.appearance_per_ray > 1 || !(d.radius > 0.0f) || !(d.radius < 1.0e18f)) {
        set_error(kBadDescriptor,
                  "%s: invalid descriptor (n_extra_dims=%u, at most %u; appearance_per_ray=%u, 0 or 1; radius=%g, "
                  "finite and positive)",
                  target, d.n_extra_dims, (unsigned)kMaxE, d.appearance_per_ray, (double)d.radius);
        return false;
    }
    return true;
}

size_t workspace_bytes(const SkySphereBgDescriptor &d) {
    const size_t b = kRowsOffset + (size_t)bwd_blocks(d.n_rays) * n_params(d.n_extra_dims) * sizeof(float);
    return (b + 255) & ~(size_t)255;
}

SkyArgs make_args(const SkySphereBgDescriptor &d) {
    const double r = d.radius;
    return SkyArgs{d.n_rays, d.n_extra_dims, d.appearance_per_ray, (float)(r * r)};
}

bool aligned(const void *p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }

// Launches the two norm passes; returns the device address of D.
const float *launch_norm(const SkyArgs &a, const float *o, const float *d, uint8_t *ws, hipStream_t s) {
    double *partial = reinterpret_cast<double *>(ws);
    float *denom = reinterpret_cast<float *>(ws + kDenomOffset);
    hipLaunchKernelGGL(sky_norm_partial_kernel, dim3(kNormBlocks), dim3(kBlock), 0, s, a, o, d, partial);
    hipLaunchKernelGGL(sky_norm_final_kernel, dim3(1), dim3(kBlock), 0, s, partial, denom);
    return denom;
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" uint32_t sky_sphere_bg_n_params(uint32_t n_extra_dims) { return n_params(n_extra_dims); }

extern "C" size_t sky_sphere_bg_workspace_size(const char *opaque, size_t opaque_len) {
    const char *target = "sky_sphere_bg_workspace_size";
    SkySphereBgDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target) || !valid(d, target)) return 0;
    return workspace_bytes(d);
}

extern "C" void sky_sphere_bg(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "sky_sphere_bg";
    SkySphereBgDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target) || !valid(d, target)) return;
    if (d.n_rays == 0) return;
    if (!check_buffer_list(buffers, {0, 1, 3, 4, 5}, target)) return;
    if (d.n_extra_dims > 0 && !check_buffer_list(buffers, {2}, target)) return;
    if (!aligned(buffers[3], 16) || !aligned(buffers[5], 16)) {
        set_error(kBadArgument, "%s: params and workspace must be 16-byte aligned", target);
        return;
    }
    const SkyArgs a = make_args(d);
    hipStream_t s = (hipStream_t)stream;
    const float *o = (const float *)buffers[0];
    const float *dv = (const float *)buffers[1];
    const float *denom = launch_norm(a, o, dv, (uint8_t *)buffers[5], s);
    const uint32_t grid = std::min<uint32_t>(div_up(d.n_rays, kBlock), 2048);
    hipLaunchKernelGGL(sky_forward_kernel, dim3(grid), dim3(kBlock), 0, s, a, o, dv, (const float *)buffers[2],
                       (const float *)buffers[3], denom, (float *)buffers[4]);
    check_launch(target);
}

extern "C" void sky_sphere_bg_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "sky_sphere_bg_backward";
    SkySphereBgDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target) || !valid(d, target)) return;
    if (d.n_extra_dims > 0 && !d.appearance_per_ray) {
        set_error(kBadDescriptor, "%s: the backward needs per-ray appearance embeddings when n_extra_dims > 0",
                  target);
        return;
    }
    if (!check_buffer_list(buffers, {5, 7}, target)) return;
    if (d.n_rays > 0 && !check_buffer_list(buffers, {0, 1, 3, 4}, target)) return;
    if (d.n_rays > 0 && d.n_extra_dims > 0 && !check_buffer_list(buffers, {2, 6}, target)) return;
    if (!aligned(buffers[3], 16) || !aligned(buffers[7], 16)) {
        set_error(kBadArgument, "%s: params and workspace must be 16-byte aligned", target);
        return;
    }
    const SkyArgs a = make_args(d);
    const uint32_t P = n_params(d.n_extra_dims);
    hipStream_t s = (hipStream_t)stream;
    uint8_t *ws = (uint8_t *)buffers[7];
    float *rows = reinterpret_cast<float *>(ws + kRowsOffset);
    uint32_t n_rows = 0;
    if (d.n_rays > 0) {
        const float *o = (const float *)buffers[0];
        const float *dv = (const float *)buffers[1];
        const float *denom = launch_norm(a, o, dv, ws, s);
        n_rows = bwd_blocks(d.n_rays);
        hipLaunchKernelGGL(sky_backward_kernel, dim3(n_rows), dim3(kBlock), 0, s, a, o, dv,
                           (const float *)buffers[2], (const float *)buffers[3], denom, (const float *)buffers[4],
                           (float *)buffers[6], rows);
    }
    hipLaunchKernelGGL(sky_sum_rows_kernel, dim3(div_up(P, kBlock)), dim3(kBlock), 0, s, n_rows, P, rows,
                       (float *)buffers[5]);
    check_launch(target);
}
