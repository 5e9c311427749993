// hashgrid.hip — multiresolution hash encoding, forward and table gradient.
//
// Forward restates models/encoders.py:82-233 (HashGridEncoder.__call__): the
// reference evaluates it as XLA gather + lerp fusions and materialises [L,N,8]
// index and weight arrays; here each thread computes one (point, level) pair in
// registers: normalise, scale, floor, 8 vertex indices (dense x + y res + z res^2
// below first_hash_level, else x ^ y*2654435761 ^ z*805459861, both mod T, plus the
// level's row offset), trilinear weights, weighted sum of the gathered rows.
// Threads are laid out level-fastest, so the L*F outputs of a point are written by
// L consecutive lanes as one contiguous run.
//
// Backward is the VJP of the gather (encoders.py:226) and of the weighted sum
// (encoders.py:231): dtable[idx] += w * g, with hardware f32 atomics.  Before the
// atomic, lanes of the same level whose points fall on the same table row (the
// common case on coarse levels, where consecutive samples of a ray share a cell)
// pre-add their contributions across the wave, so one atomic carries them.
//
// Arithmetic: f32, no contraction (matches the CPU oracle bit for bit on indices
// and weights).  Indices past the last row (only possible when the last level is
// dense) are clamped to the last row, as JAX clamps gather indices.
#include <hip/hip_fp16.h>

#include "ngp_common.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;

struct LevelTables {
    float scale[NGP_HASHGRID_MAX_LEVELS];
    uint32_t res[NGP_HASHGRID_MAX_LEVELS];
    uint32_t offset[NGP_HASHGRID_MAX_LEVELS];
};

template <typename TT>
__device__ __forceinline__ float to_f32(TT v);
template <>
__device__ __forceinline__ float to_f32<float>(float v) { return v; }
template <>
__device__ __forceinline__ float to_f32<__half>(__half v) { return __half2float(v); }

// Vertex rows and weights of point (px,py,pz) (already normalised to [0,1]) at one level.
__device__ __forceinline__ void level_vertices(float px, float py, float pz, float scale, uint32_t res, bool dense,
                                               uint32_t T, bool pow2, uint32_t offset, uint32_t last_row,
                                               uint32_t idx[8], float w[8]) {
    const float sx = px * scale + 0.5f, sy = py * scale + 0.5f, sz = pz * scale + 0.5f;
    const float fx = floorf(sx), fy = floorf(sy), fz = floorf(sz);
    const float rx = sx - fx, ry = sy - fy, rz = sz - fz;
    const uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy, z0 = (uint32_t)fz;
    const uint32_t x1 = (uint32_t)(fx + 1.0f), y1 = (uint32_t)(fy + 1.0f), z1 = (uint32_t)(fz + 1.0f);
    const float wx[2] = {1.0f - rx, rx}, wy[2] = {1.0f - ry, ry}, wz[2] = {1.0f - rz, rz};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ox = (c >> 2) & 1, oy = (c >> 1) & 1, oz = c & 1;
        const uint32_t vx = ox ? x1 : x0, vy = oy ? y1 : y0, vz = oz ? z1 : z0;
        uint32_t h = dense ? vx + vy * res + vz * (res * res) : vx ^ (vy * 2654435761u) ^ (vz * 805459861u);
        h = (pow2 ? (h & (T - 1u)) : (h % T)) + offset;
        idx[c] = h > last_row ? last_row : h;
        w[c] = (wx[ox] * wy[oy]) * wz[oz];
    }
}

template <int F, typename TT>
__global__ void __launch_bounds__(kBlock)
    hashgrid_fwd_kernel(HashGridDescriptor d, const float *__restrict__ pos, const TT *__restrict__ table,
                        float *__restrict__ enc) {
    __shared__ LevelTables lt;
    if (threadIdx.x < d.L) {
        lt.scale[threadIdx.x] = d.scales[threadIdx.x];
        lt.res[threadIdx.x] = d.resolutions[threadIdx.x];
        lt.offset[threadIdx.x] = d.offsets[threadIdx.x];
    }
    __syncthreads();
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= (uint64_t)d.n_points * d.L) return;
    const uint32_t n = (uint32_t)(gid / d.L);
    const uint32_t l = (uint32_t)(gid - (uint64_t)n * d.L);
    const float two_b = 2.0f * d.bound;
    const float px = (pos[3 * (size_t)n + 0] + d.bound) / two_b;
    const float py = (pos[3 * (size_t)n + 1] + d.bound) / two_b;
    const float pz = (pos[3 * (size_t)n + 2] + d.bound) / two_b;
    uint32_t idx[8];
    float w[8];
    level_vertices(px, py, pz, lt.scale[l], lt.res[l], l < d.first_hash_level, d.T, (d.T & (d.T - 1u)) == 0,
                   lt.offset[l], d.n_rows - 1u, idx, w);
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const TT *row = table + (size_t)idx[c] * F;
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = acc[f] + to_f32<TT>(row[f]) * w[c];
    }
    float *out = enc + gid * F;
#pragma unroll
    for (int f = 0; f < F; ++f) out[f] = acc[f];
}

template <int F>
__global__ void __launch_bounds__(kBlock)
    hashgrid_bwd_kernel(HashGridDescriptor d, const float *__restrict__ pos, const float *__restrict__ dL_denc,
                        float *__restrict__ dtable) {
    __shared__ LevelTables lt;
    if (threadIdx.x < d.L) {
        lt.scale[threadIdx.x] = d.scales[threadIdx.x];
        lt.res[threadIdx.x] = d.resolutions[threadIdx.x];
        lt.offset[threadIdx.x] = d.offsets[threadIdx.x];
    }
    __syncthreads();
    const uint64_t total = (uint64_t)d.n_points * d.L;
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = gid < total;
    const uint32_t n = active ? (uint32_t)(gid / d.L) : 0u;
    const uint32_t l = active ? (uint32_t)(gid - (uint64_t)n * d.L) : 0u;
    const float two_b = 2.0f * d.bound;
    const float px = (pos[3 * (size_t)n + 0] + d.bound) / two_b;
    const float py = (pos[3 * (size_t)n + 1] + d.bound) / two_b;
    const float pz = (pos[3 * (size_t)n + 2] + d.bound) / two_b;
    uint32_t idx[8];
    float w[8];
    level_vertices(px, py, pz, lt.scale[l], lt.res[l], l < d.first_hash_level, d.T, (d.T & (d.T - 1u)) == 0,
                   lt.offset[l], d.n_rows - 1u, idx, w);
    float g[F];
#pragma unroll
    for (int f = 0; f < F; ++f) g[f] = active ? dL_denc[gid * F + f] : 0.0f;
    // Lanes l, l+L, l+2L, ... of a wave hold the same level for consecutive points.
    // When L divides 64, fold lane pairs at distances 32, 16, ... >= L that carry
    // the same row: the lower lane adds the upper lane's contribution.
    const int lane = lane_id();
    const bool fold = active && (kWave % d.L) == 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float v[F];
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] = w[c] * g[f];
        bool owner = active;
        uint32_t my = idx[c];
        for (int dist = kWave / 2; dist >= (int)d.L; dist >>= 1) {
            // every lane takes part in the shuffles; `fold` decides whether to merge
            const uint32_t other_idx = __shfl_xor(my, dist, kWave);
            const int other_owner = __shfl_xor(owner ? 1 : 0, dist, kWave);
            float ov[F];
#pragma unroll
            for (int f = 0; f < F; ++f) ov[f] = __shfl_xor(v[f], dist, kWave);
            if (fold && other_idx == my && other_owner && owner) {
                if ((lane & dist) == 0) {
#pragma unroll
                    for (int f = 0; f < F; ++f) v[f] += ov[f];
                } else {
                    owner = false;
                }
            }
        }
        if (owner) {
            float *dst = dtable + (size_t)my * F;
#pragma unroll
            for (int f = 0; f < F; ++f)
                if (v[f] != 0.0f) unsafeAtomicAdd(dst + f, v[f]);
        }
    }
}

__global__ void zero_table_kernel(float4 *__restrict__ a, size_t n4, float *__restrict__ tail, size_t n_tail) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t j = g; j < n4; j += stride) a[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (g < n_tail) tail[g] = 0.0f;
}

bool validate(const HashGridDescriptor &d, const char *target) {
    if (d.L == 0 || d.L > NGP_HASHGRID_MAX_LEVELS) {
        set_error(kBadDescriptor, "%s: L=%u must be in [1, %d]", target, d.L, NGP_HASHGRID_MAX_LEVELS);
        return false;
    }
    if (!(d.F == 1 || d.F == 2 || d.F == 4 || d.F == 8)) {
        set_error(kBadDescriptor, "%s: F=%u must be 1, 2, 4 or 8", target, d.F);
        return false;
    }
    if (d.T == 0 || d.first_hash_level > d.L || d.n_rows == 0 || d.offsets[0] != 0 || d.offsets[d.L] != d.n_rows ||
        !(d.bound > 0.0f)) {
        set_error(kBadDescriptor, "%s: inconsistent level table (T=%u first_hash_level=%u n_rows=%u bound=%g)", target,
                  d.T, d.first_hash_level, d.n_rows, (double)d.bound);
        return false;
    }
    for (uint32_t l = 0; l < d.L; ++l) {
        if (d.offsets[l + 1] < d.offsets[l]) {
            set_error(kBadDescriptor, "%s: offsets must be non-decreasing", target);
            return false;
        }
    }
    if (d.table_dtype != NGP_DTYPE_F32 && d.table_dtype != NGP_DTYPE_F16) {
        set_error(kBadDescriptor, "%s: table_dtype=%u must be 0 (f32) or 1 (f16)", target, d.table_dtype);
        return false;
    }
    return true;
}

template <typename TT>
void launch_fwd(const HashGridDescriptor &d, hipStream_t s, const float *pos, const void *table, float *enc) {
    const uint32_t grid = div_up((uint64_t)d.n_points * d.L, kBlock);
    const TT *t = static_cast<const TT *>(table);
    switch (d.F) {
        case 1: hipLaunchKernelGGL((hashgrid_fwd_kernel<1, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
        case 2: hipLaunchKernelGGL((hashgrid_fwd_kernel<2, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
        case 4: hipLaunchKernelGGL((hashgrid_fwd_kernel<4, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
        default: hipLaunchKernelGGL((hashgrid_fwd_kernel<8, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
    }
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void hashgrid_encode(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "hashgrid_encode";
    HashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!validate(d, target)) return;
    if (d.n_points == 0) return;
    if (!check_buffers(buffers, 3, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    const float *pos = (const float *)buffers[0];
    float *enc = (float *)buffers[2];
    if (d.table_dtype == NGP_DTYPE_F32) launch_fwd<float>(d, stream, pos, buffers[1], enc);
    else launch_fwd<__half>(d, stream, pos, buffers[1], enc);
    check_launch(target);
}

extern "C" void hashgrid_encode_backward(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "hashgrid_encode_backward";
    HashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!validate(d, target)) return;
    if (!check_buffers(buffers, 3, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    const float *pos = (const float *)buffers[0];
    const float *g = (const float *)buffers[1];
    float *dtable = (float *)buffers[2];
    const size_t n = (size_t)d.n_rows * d.F;
    const bool vec = (reinterpret_cast<uintptr_t>(dtable) & 15u) == 0;
    const size_t n4 = vec ? n / 4 : 0;
    const size_t n_tail = n - 4 * n4;
    const uint32_t zgrid = (uint32_t)std::max<size_t>(1, std::min<size_t>((std::max(n4, n_tail) + 255) / 256, 16384));
    hipLaunchKernelGGL(zero_table_kernel, dim3(zgrid), dim3(256), 0, stream, reinterpret_cast<float4 *>(dtable), n4,
                       dtable + 4 * n4, n_tail);
    if (!check_launch(target) || d.n_points == 0) return;
    const uint32_t grid = div_up((uint64_t)d.n_points * d.L, kBlock);
    switch (d.F) {
        case 1: hipLaunchKernelGGL(hashgrid_bwd_kernel<1>, dim3(grid), dim3(kBlock), 0, stream, d, pos, g, dtable); break;
        case 2: hipLaunchKernelGGL(hashgrid_bwd_kernel<2>, dim3(grid), dim3(kBlock), 0, stream, d, pos, g, dtable); break;
        case 4: hipLaunchKernelGGL(hashgrid_bwd_kernel<4>, dim3(grid), dim3(kBlock), 0, stream, d, pos, g, dtable); break;
        default: hipLaunchKernelGGL(hashgrid_bwd_kernel<8>, dim3(grid), dim3(kBlock), 0, stream, d, pos, g, dtable); break;
    }
    check_launch(target);
}
