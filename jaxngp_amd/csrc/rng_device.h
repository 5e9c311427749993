// rng_device.h — the stateless counter-based random numbers of the training step.
//
// A draw is a 64-bit mix of (seed, step, item, stream), so every random number a
// kernel uses is reproducible from its inputs and no generator state lives on the
// device.  The mixer is the splitmix64 finaliser (Steele, Lea, Flood 2014).
#pragma once

#include <cstdint>

namespace ngp {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t draw(uint64_t seed, uint64_t step, uint64_t i, uint64_t stream) {
    return mix64(seed ^ mix64(step * 0x9e3779b97f4a7c15ull ^ mix64(i * 4 + stream)));
}

// uniform in [0, 1) with 24 random bits
__device__ __forceinline__ float unit24(uint64_t r) { return (float)(uint32_t)(r >> 40) * (1.0f / 16777216.0f); }

// an index in [0, n): the top 32 bits of the draw scaled by n (multiply-shift)
__device__ __forceinline__ uint32_t below(uint64_t r, uint32_t n) {
    return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

}  // namespace ngp
