// ngp_common.h — shared host/device helpers of libngp_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <initializer_list>

#include "ngp_hip.h"

namespace ngp {

// ---------------------------------------------------------------------------
// error reporting (status.cpp).  Entry points never throw across the C ABI.
// ---------------------------------------------------------------------------
enum ErrorCode : int {
    kOk = 0,
    kBadDescriptor = 1,
    kBadArgument = 2,
    kLaunchFailed = 3,
};

void set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
bool has_error();

// Copies the opaque bytes into `out` after checking their length
// (the reference's deserialize, serde.h:35-41, throws on a mismatch instead).
template <typename T>
inline bool read_descriptor(const char *opaque, size_t opaque_len, T *out, const char *target) {
    if (opaque == nullptr || opaque_len != sizeof(T)) {
        set_error(kBadDescriptor, "%s: invalid opaque descriptor size, expected %zu, got %zu", target, sizeof(T),
                  opaque_len);
        return false;
    }
    std::memcpy(out, opaque, sizeof(T));
    return true;
}

inline bool check_buffers(void **buffers, int n, const char *target) {
    if (buffers == nullptr) {
        set_error(kBadArgument, "%s: buffers is NULL", target);
        return false;
    }
    for (int i = 0; i < n; ++i) {
        if (buffers[i] == nullptr) {
            set_error(kBadArgument, "%s: buffer %d is NULL", target, i);
            return false;
        }
    }
    return true;
}

// Checks only the listed buffers (zero-sized operands may legitimately be NULL).
inline bool check_buffer_list(void **buffers, std::initializer_list<int> which, const char *target) {
    if (buffers == nullptr) {
        set_error(kBadArgument, "%s: buffers is NULL", target);
        return false;
    }
    for (int i : which) {
        if (buffers[i] == nullptr) {
            set_error(kBadArgument, "%s: buffer %d is NULL", target, i);
            return false;
        }
    }
    return true;
}

inline bool check_launch(const char *target) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(kLaunchFailed, "%s: kernel launch failed: %s", target, hipGetErrorString(e));
        return false;
    }
    return true;
}

inline uint32_t div_up(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// device math
// ---------------------------------------------------------------------------

// One element of optax.adam(lr, b1, b2, eps, eps_root) followed by
// add_decayed_weights(weight_decay), in optax's order of operations
// (app/nerf/_utils.py:19-77); gradients are multiplied by gs first.  Shared by
// adam_update and the table gradient's fused update, so both round alike.
struct AdamStep {
    float lr, b1, b2, eps, eps_root, c1, c2, weight_decay, gs;
};

__host__ __device__ inline AdamStep make_adam_step(const NgpAdamDescriptor &d) {
    return AdamStep{d.lr, d.b1, d.b2, d.eps, d.eps_root, 1.0f - d.b1_pow_t, 1.0f - d.b2_pow_t, d.weight_decay,
                    d.grad_scale};
}

__device__ __forceinline__ float adam_step(const AdamStep &a, float pk, float gk_raw, float &mk, float &vk) {
    const float gk = gk_raw * a.gs;
    mk = a.b1 * mk + (1.0f - a.b1) * gk;
    vk = a.b2 * vk + (1.0f - a.b2) * (gk * gk);
    const float mh = mk / a.c1, vh = vk / a.c2;
    const float u = -a.lr * (mh / (sqrtf(vh + a.eps_root) + a.eps)) + a.weight_decay * pk;
    return pk + u;
}

// x / d for a divisor d shared by a whole launch: a power of two in the normal range
// takes the multiply by its exact reciprocal (x / 2^k and x * 2^-k are the same exact
// scaling, rounded alike, also below the normal range), any other d the division.
struct Divisor {
    float d, inv;
    bool pow2;
};

__host__ __device__ inline Divisor make_divisor(float d) {
    Divisor q;
    q.d = d;
    uint32_t b;
    __builtin_memcpy(&b, &d, sizeof b);
    q.pow2 = (b & 0x007fffffu) == 0u && d >= 1.1754944e-38f && d <= 8.5e37f;
    const uint32_t ib = (254u << 23) - b;
    __builtin_memcpy(&q.inv, &ib, sizeof ib);
    if (!q.pow2) q.inv = 0.0f;
    return q;
}

__device__ __forceinline__ float divide(const Divisor &q, float x) { return q.pow2 ? x * q.inv : x / q.d; }

// Deterministic exp: exp(x) = 2^k exp(r), k = rint(x/ln2), r = x - k ln2 by the
// two-constant reduction of Cody & Waite, exp(r) by its degree-7 Taylor polynomial
// in Horner form with fused multiply-adds, 2^k applied as two exact power-of-two
// factors.  Built only from correctly rounded operations, so the CPU oracle
// reproduces it bit for bit.
__device__ __forceinline__ float det_expf(float x) {
    if (x != x) return x;
    if (x < -104.0f) return 0.0f;
    if (x > 88.72283935546875f) return __builtin_huge_valf();
    float k = __builtin_rintf(x * 1.44269502162933349609375f);
    float r = __builtin_fmaf(k, -0.693145751953125f, x);
    r = __builtin_fmaf(k, -1.428606765330187045e-06f, r);
    float p = 1.98412701e-04f;
    p = __builtin_fmaf(p, r, 1.38888892e-03f);
    p = __builtin_fmaf(p, r, 8.33333377e-03f);
    p = __builtin_fmaf(p, r, 4.16666679e-02f);
    p = __builtin_fmaf(p, r, 1.66666672e-01f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    int ki = (int)k;
    int k1 = ki / 2;
    int k2 = ki - k1;
    float s1 = __uint_as_float((uint32_t)(k1 + 127) << 23);
    float s2 = __uint_as_float((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// ---------------------------------------------------------------------------
// wavefront (64 lanes) primitives
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Inclusive prefix sum over the 64 lanes of the wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Float sum over the wavefront (butterfly: every lane gets the same value).
__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Exclusive prefix sum over a block of BLOCK threads (BLOCK a multiple of 64).
// Returns the exclusive prefix of `v`; `total` receives the block sum.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    constexpr int kWaves = BLOCK / kWave;
    __shared__ uint32_t wave_totals[kWaves];
    const int wave = (int)threadIdx.x / kWave;
    const int lane = lane_id();
    uint32_t incl = wave_inclusive_scan(v);
    if (lane == kWave - 1) wave_totals[wave] = incl;
    __syncthreads();
    uint32_t base = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        uint32_t t = wave_totals[w];
        if (w < wave) base += t;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return base + incl - v;
}

}  // namespace ngp
