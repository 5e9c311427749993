// hashgrid_device.h — per-level vertex rows and weights of the hash encoding
// (models/encoders.py:116-213), shared by the encoder kernels and the fused
// inference field kernel.
#pragma once

#include <hip/hip_fp16.h>

#include "ngp_common.h"

namespace ngp {

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
template <>
__device__ __forceinline__ float to_f32<_Float16>(_Float16 v) { return (float)v; }

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

}  // namespace ngp
