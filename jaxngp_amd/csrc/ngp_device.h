// ngp_device.h — device helpers of the occupancy-grid marcher.
//
// Arithmetic follows one convention shared with the CPU oracle: f32, every
// operation rounded on its own (sources built with -ffp-contract=off), IEEE
// division.  Cited reference lines are in
// deps/volume-rendering-jax/lib/impl/marching.cu.
#pragma once

#include "ngp_common.h"

namespace ngp {

constexpr float kSqrt3 = 1.7320508075688772f;  // volrend.h:18
constexpr float kTThreshold = 1e-4f;           // integrating.cu:12

// 10-bit bit interleave by multiply-and-mask (marching.cu:52-68).
__device__ __forceinline__ uint32_t morton_spread(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return morton_spread(x) | (morton_spread(y) << 1) | (morton_spread(z) << 2);
}
// Keep every third bit and compact (marching.cu:70-77).
__device__ __forceinline__ uint32_t morton_compact(uint32_t x) {
    x = x & 0x49249249u;
    x = (x | (x >> 2)) & 0xc30c30c3u;
    x = (x | (x >> 4)) & 0x0f00f00fu;
    x = (x | (x >> 8)) & 0xff0000ffu;
    x = (x | (x >> 16)) & 0x0000ffffu;
    return x;
}

struct MarchParams {
    uint32_t diag, K, G, G3;
    float bound, portion, inv_G, ds_lo, ds_hi, sample_limit;
    // samples a ray may take: the loop test (float)n < sample_limit holds exactly
    // for n < n_cap
    uint32_t n_cap;
};

inline MarchParams make_march_params(uint32_t diag, uint32_t K, uint32_t G, float bound, float portion) {
    MarchParams p;
    p.diag = diag;
    p.K = K;
    p.G = G;
    p.G3 = G * G * G;
    p.bound = bound;
    p.portion = portion;
    p.inv_G = 1.0f / (float)G;
    // marching.cu:18-22
    p.ds_lo = ((2.0f * kSqrt3) * (bound < 1.0f ? bound : 1.0f)) / (float)diag;
    p.ds_hi = ((2.0f * kSqrt3) * bound) * p.inv_G;
    p.sample_limit = (float)diag * bound;  // marching.cu:165
    // smallest n with (float)n >= sample_limit ((float)n is monotone in n)
    uint64_t lo = 0, hi = 0x100000000ull;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if ((float)mid >= p.sample_limit) hi = mid;
        else lo = mid + 1;
    }
    p.n_cap = lo > 0xffffffffull ? 0xffffffffu : (uint32_t)lo;
    return p;
}

// marching.cu:15-23
__device__ __forceinline__ float step_size(const MarchParams &p, float t) {
    return fminf(fmaxf(t * p.portion, p.ds_lo), p.ds_hi);
}

__device__ __forceinline__ int frexp_exp(float x) {
    int e;
    frexpf(x, &e);
    return e;
}

__device__ __forceinline__ uint32_t clamp_cascade(int e, uint32_t K) {
    int hi = (int)K - 1;
    return (uint32_t)(e < 0 ? 0 : (e > hi ? hi : e));
}

// The last 64-bit word of the occupancy bitfield a ray read.  A word holds 64
// Morton-consecutive cells (a 4x4x4 block); consecutive steps of a ray mostly stay
// inside one, so most lookups need no memory access.  `words` is NULL when the
// bitfield is too small or not 8-byte aligned; then single bytes are read.
struct BitCache {
    const uint64_t *__restrict__ words;
    uint64_t word;
    uint32_t key;
};

__device__ __forceinline__ BitCache make_bit_cache(const uint8_t *bits, bool word_ok) {
    BitCache c;
    c.words = word_ok ? reinterpret_cast<const uint64_t *>(bits) : nullptr;
    c.word = 0;
    c.key = 0xffffffffu;
    return c;
}

__device__ __forceinline__ bool occupied(BitCache &c, const uint8_t *__restrict__ bits, uint32_t idx) {
    if (c.words == nullptr) return (bits[idx >> 3] >> (idx & 7u)) & 1u;
    const uint32_t key = idx >> 6;
    if (key != c.key) {
        c.word = c.words[key];
        c.key = key;
    }
    return (c.word >> (idx & 63u)) & 1u;
}

// Occupancy cell of one position (marching.cu:79-98 and 170-176): its index in the
// bitfield, the float grid coordinate and the mip bound used by the empty-cell skip.
__device__ __forceinline__ uint32_t cell_index(const MarchParams &p, float px, float py, float pz, float ds, float &gx,
                                               float &gy, float &gz, float &mip_bound) {
    uint32_t cascade = 0;
    if (p.K > 1) {
        float m = fmaxf(fabsf(px), fmaxf(fabsf(py), fabsf(pz)));
        uint32_t c1 = clamp_cascade(frexp_exp(m), p.K);
        uint32_t c2 = clamp_cascade(frexp_exp(ds * (float)p.G), p.K);
        cascade = c1 > c2 ? c1 : c2;
    }
    mip_bound = fminf(__uint_as_float((cascade + 127u) << 23), p.bound);
    // x / 2^k and x * 2^-k are the same exact scaling, rounded alike (also below the
    // normal range), so a power-of-two bound (every cascade but a clamped last one)
    // takes a multiply instead of the division
    float qx, qy, qz;
    if ((__float_as_uint(mip_bound) & 0x007fffffu) == 0u && mip_bound >= 1.1754944e-38f && mip_bound <= 8.5e37f) {
        const float inv = __uint_as_float((254u << 23) - __float_as_uint(mip_bound));  // 1 / mip_bound, exact
        qx = px * inv;
        qy = py * inv;
        qz = pz * inv;
    } else {
        qx = px / mip_bound;
        qy = py / mip_bound;
        qz = pz / mip_bound;
    }
    const float G = (float)p.G;
    gx = ((qx + 1.0f) * 0.5f) * G;
    gy = ((qy + 1.0f) * 0.5f) * G;
    gz = ((qz + 1.0f) * 0.5f) * G;
    const float gmax = (float)(p.G - 1);
    // clamp in float: identical to clampi((int)floorf(.), 0, G-1) for every input
    uint32_t ix = (uint32_t)fminf(fmaxf(floorf(gx), 0.0f), gmax);
    uint32_t iy = (uint32_t)fminf(fmaxf(floorf(gy), 0.0f), gmax);
    uint32_t iz = (uint32_t)fminf(fmaxf(floorf(gz), 0.0f), gmax);
    return cascade * p.G3 + morton3(ix, iy, iz);
}

// Occupancy lookup of one position through the ray's bit cache.
__device__ __forceinline__ bool lookup_cell(const MarchParams &p, const uint8_t *__restrict__ bits, BitCache &cache,
                                            float px, float py, float pz, float ds, float &gx, float &gy, float &gz,
                                            float &mip_bound) {
    return occupied(cache, bits, cell_index(p, px, py, pz, ds, gx, gy, gz, mip_bound));
}

// Exit distance of the current cell along the ray (marching.cu:182-185).
__device__ __forceinline__ float cell_exit_t(const MarchParams &p, float t, float px, float py, float pz, float gx,
                                             float gy, float gz, float mip_bound, float dx, float dy, float dz,
                                             float idx_, float idy, float idz) {
    float nx = floorf((gx + 0.5f) + 0.5f * copysignf(1.0f, dx));
    float ny = floorf((gy + 0.5f) + 0.5f * copysignf(1.0f, dy));
    float nz = floorf((gz + 0.5f) + 0.5f * copysignf(1.0f, dz));
    float ex = ((((nx * p.inv_G) - 0.5f) * 2.0f) * mip_bound - px) * idx_;
    float ey = ((((ny * p.inv_G) - 0.5f) * 2.0f) * mip_bound - py) * idy;
    float ez = ((((nz * p.inv_G) - 0.5f) * 2.0f) * mip_bound - pz) * idz;
    return t + fmaxf(0.0f, fminf(ex, fminf(ey, ez)));
}

// Steps by ds from t until past next_t (marching.cu:186-189).  A step that no
// longer changes t (where the reference would loop forever) ends the march:
// returns +inf, so every loop here terminates.
__device__ __forceinline__ float skip_to(const MarchParams &p, float t, float ds, float next_t) {
    float nt = t + ds;
    if (nt == t) return __builtin_huge_valf();
    while (nt < next_t) {
        float s = nt + step_size(p, nt);
        if (s == nt) return __builtin_huge_valf();
        nt = s;
    }
    return nt;
}

}  // namespace ngp
