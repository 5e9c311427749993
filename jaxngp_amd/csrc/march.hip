// march.hip — occupancy-grid ray marching for training and inference.
//
// march_rays:            deps/volume-rendering-jax/lib/impl/marching.cu:101-268 (kernel), 435-522 (launcher)
// march_rays_inference:  marching.cu:271-397 (kernel), 524-604 (launcher)
//
// Training marcher.  The reference reserves each ray's output range with an
// atomicAdd in scheduling order and marches every ray twice.  Here the ranges are
// an exclusive prefix sum of the per-ray counts in ray order, which makes the
// output deterministic and equal to processing rays one after another:
//   1. march_count:  one thread per ray marches and counts occupied steps; the
//                    block total is parked in rays_sample_startidx[first ray of
//                    the block] (slot owned by the block, so no scratch memory).
//                    Rays are counted in three ray-ordered launches (8192 rays,
//                    up to 65536, the rest); a launch whose predecessors already
//                    filled the sample buffer skips marching, which changes no
//                    output (those rays are invalid regardless of their counts);
//   2. march_scan:   one workgroup scans the block totals, finds the ray where the
//                    sample buffer fills up and writes the two helper counters;
//   3. march_write:  one thread per ray takes its offset (block prefix + wave/LDS
//                    scan), decides validity, re-marches valid rays writing their
//                    samples, and zero-fills the unused tail of the sample arrays.
// A ray is valid iff the samples of the rays before it leave room (prefix <
// total_samples), t_end > t_start, and its own samples fit; rays after the buffer
// fills are invalid, as the reference's early exit at marching.cu:135 makes them.
#include "ngp_common.h"
#include "ngp_device.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;
constexpr int kScanThreads = 1024;

// Counts (or, with `write`, emits) the samples of one ray (marching.cu:162-192, 224-267).
template <bool kWrite>
__device__ uint32_t march_ray(const MarchParams &p, const uint8_t *__restrict__ bits, bool word_ok, float ox, float oy,
                              float oz,
                              float dx, float dy, float dz, float t0, float t1, float noise, uint32_t max_samples,
                              uint32_t ray, uint32_t *__restrict__ idcs, float *__restrict__ xyzs,
                              float *__restrict__ dirs, float *__restrict__ dss, float *__restrict__ zs) {
    const float idx_ = 1.0f / dx, idy = 1.0f / dy, idz = 1.0f / dz;
    BitCache cache = make_bit_cache(bits, word_ok);
    float t = t0;
    t = t + step_size(p, t) * noise;
    uint32_t n = 0;
    while ((float)n < p.sample_limit && n < max_samples && t < t1) {
        const float px = ox + t * dx, py = oy + t * dy, pz = oz + t * dz;
        const float ds = step_size(p, t);
        float gx, gy, gz, mb;
        const bool occ = lookup_cell(p, bits, cache, px, py, pz, ds, gx, gy, gz, mb);
        float nt;
        if (occ) {
            if (kWrite) {
                idcs[n] = ray;
                xyzs[3 * n + 0] = px;
                xyzs[3 * n + 1] = py;
                xyzs[3 * n + 2] = pz;
                dirs[3 * n + 0] = dx;
                dirs[3 * n + 1] = dy;
                dirs[3 * n + 2] = dz;
                dss[n] = ds;
                zs[n] = t;
            }
            ++n;
            nt = t + ds;
        } else {
            const float next_t = cell_exit_t(p, t, px, py, pz, gx, gy, gz, mb, dx, dy, dz, idx_, idy, idz);
            nt = skip_to(p, t, ds, next_t);
        }
        t = nt;
    }
    return n;
}

// Counts rays of blocks [block_begin, block_end).  Blocks of later chunks first
// add up the parked totals of all earlier blocks (final: earlier chunks ran to
// completion before this launch); when they already reach total_samples, every ray
// here starts past the filled buffer and is invalid whatever its count, so the
// block records zero samples without marching.
__global__ void __launch_bounds__(kBlock)
    march_count_kernel(MarchParams p, bool word_ok, uint32_t n_rays, uint32_t total_samples, uint32_t block_begin,
                       const float *__restrict__ rays_o, const float *__restrict__ rays_d,
                       const float *__restrict__ t_starts, const float *__restrict__ t_ends,
                       const float *__restrict__ noises, const uint8_t *__restrict__ bits,
                       uint32_t *__restrict__ rays_n_samples,
                       uint32_t *__restrict__ block_totals /* = rays_sample_startidx */) {
    __shared__ uint32_t wave_totals[kBlock / kWave];
    __shared__ uint32_t skip;
    const uint32_t blk = block_begin + blockIdx.x;
    if (threadIdx.x < kWave) {
        uint64_t prior = 0;
        for (uint32_t b = threadIdx.x; b < block_begin; b += kWave) prior += block_totals[(size_t)b * kBlock];
        prior = wave_sum64(prior);
        if (threadIdx.x == 0) skip = prior >= total_samples ? 1u : 0u;
    }
    __syncthreads();
    const uint32_t i = blk * kBlock + threadIdx.x;
    uint32_t n = 0;
    if (i < n_rays) {
        const float t0 = t_starts[i], t1 = t_ends[i];
        if (!skip && t1 > t0) {
            n = march_ray<false>(p, bits, word_ok, rays_o[3 * (size_t)i], rays_o[3 * (size_t)i + 1], rays_o[3 * (size_t)i + 2],
                                 rays_d[3 * (size_t)i], rays_d[3 * (size_t)i + 1], rays_d[3 * (size_t)i + 2], t0, t1,
                                 noises[i], 0xffffffffu, i, nullptr, nullptr, nullptr, nullptr, nullptr);
        }
        rays_n_samples[i] = n;
    }
    const uint32_t ws = wave_sum(n);
    if (lane_id() == 0) wave_totals[threadIdx.x / kWave] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) s += wave_totals[w];
        block_totals[(size_t)blk * kBlock] = s;
    }
}

// Single workgroup.  Replaces each parked block total with the block's exclusive
// prefix (saturated to u32; only its comparison with total_samples matters) and
// writes next_sample_write_location / number_of_exceeded_samples.
__global__ void __launch_bounds__(kScanThreads)
    march_scan_kernel(uint32_t n_rays, uint32_t n_blocks, uint32_t total_samples,
                      const uint32_t *__restrict__ rays_n_samples, uint32_t *__restrict__ block_slots,
                      uint32_t *__restrict__ next_loc, uint32_t *__restrict__ exceeded) {
    __shared__ uint64_t partial[kScanThreads];
    __shared__ uint32_t fill_block;
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n_blocks + kScanThreads - 1) / kScanThreads;
    const uint32_t b0 = tid * per;
    uint64_t local = 0;
    for (uint32_t k = 0; k < per; ++k) {
        uint32_t b = b0 + k;
        if (b < n_blocks) local += block_slots[(size_t)b * kBlock];
    }
    partial[tid] = local;
    if (tid == 0) fill_block = 0xffffffffu;
    __syncthreads();
    // inclusive Hillis-Steele scan of the per-thread sums (u64)
    for (uint32_t d = 1; d < kScanThreads; d <<= 1) {
        uint64_t v = tid >= d ? partial[tid - d] : 0;
        __syncthreads();
        partial[tid] += v;
        __syncthreads();
    }
    const uint64_t grand = partial[kScanThreads - 1];
    uint64_t run = partial[tid] - local;  // exclusive prefix of this thread's first block
    for (uint32_t k = 0; k < per; ++k) {
        uint32_t b = b0 + k;
        if (b >= n_blocks) break;
        uint32_t tot = block_slots[(size_t)b * kBlock];
        // first block whose inclusive prefix passes total_samples holds the fill point
        if (run + tot > total_samples && run <= total_samples) atomicMin(&fill_block, b);
        block_slots[(size_t)b * kBlock] = run > 0xffffffffull ? 0xffffffffu : (uint32_t)run;
        run += tot;
    }
    __syncthreads();
    if (tid == 0) {
        if (grand <= total_samples) {
            *next_loc = (uint32_t)grand;
            *exceeded = 0;
        } else {
            // walk the fill block's rays: first ray whose inclusive prefix passes total_samples
            uint32_t b = fill_block < n_blocks ? fill_block : n_blocks - 1;  // always set when grand > S
            uint64_t pre = block_slots[(size_t)b * kBlock];
            uint32_t first = b * kBlock, last = min(n_rays, first + kBlock);
            uint32_t measured = (uint32_t)pre, over = 0;
            for (uint32_t r = first; r < last; ++r) {
                uint32_t n = rays_n_samples[r];
                if (pre + n > total_samples) {
                    measured = (uint32_t)pre;
                    over = pre < total_samples ? n : 0;
                    break;
                }
                pre += n;
            }
            *next_loc = measured + over;
            *exceeded = over;
        }
    }
}

__global__ void __launch_bounds__(kBlock)
    march_write_kernel(MarchParams p, bool word_ok, uint32_t n_rays, uint32_t total_samples,
                       const float *__restrict__ rays_o,
                       const float *__restrict__ rays_d, const float *__restrict__ t_starts,
                       const float *__restrict__ t_ends, const float *__restrict__ noises,
                       const uint8_t *__restrict__ bits, const uint32_t *__restrict__ next_loc,
                       const uint32_t *__restrict__ exceeded, bool *__restrict__ ray_is_valid,
                       uint32_t *__restrict__ rays_n_samples, uint32_t *__restrict__ rays_sample_startidx,
                       uint32_t *__restrict__ idcs, float *__restrict__ xyzs, float *__restrict__ dirs,
                       float *__restrict__ dss, float *__restrict__ zs) {
    __shared__ uint32_t block_prefix;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (threadIdx.x == 0) block_prefix = rays_sample_startidx[(size_t)blockIdx.x * kBlock];
    const uint32_t n = i < n_rays ? rays_n_samples[i] : 0u;
    __syncthreads();
    uint32_t block_total;
    const uint32_t local = block_exclusive_scan<kBlock>(n, &block_total);
    const uint64_t start = (uint64_t)block_prefix + local;
    if (i < n_rays) {
        const float t0 = t_starts[i], t1 = t_ends[i];
        const bool valid = start < total_samples && t1 > t0 && start + n <= total_samples;
        ray_is_valid[i] = valid;
        rays_n_samples[i] = valid ? n : 0u;
        rays_sample_startidx[i] = (valid && n > 0) ? (uint32_t)start : 0u;
        if (valid && n > 0) {
            const uint32_t s = (uint32_t)start;
            march_ray<true>(p, bits, word_ok, rays_o[3 * (size_t)i], rays_o[3 * (size_t)i + 1], rays_o[3 * (size_t)i + 2],
                            rays_d[3 * (size_t)i], rays_d[3 * (size_t)i + 1], rays_d[3 * (size_t)i + 2], t0, t1,
                            noises[i], n, i, idcs + s, xyzs + 3 * (size_t)s, dirs + 3 * (size_t)s, dss + s, zs + s);
        }
    }
    // zero the sample slots no valid ray owns: [measured, total_samples)
    const uint32_t measured = *next_loc - *exceeded;
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t j = (size_t)measured + blockIdx.x * kBlock + threadIdx.x; j < total_samples; j += stride) {
        idcs[j] = 0u;
        xyzs[3 * j + 0] = 0.0f;
        xyzs[3 * j + 1] = 0.0f;
        xyzs[3 * j + 2] = 0.0f;
        dirs[3 * j + 0] = 0.0f;
        dirs[3 * j + 1] = 0.0f;
        dirs[3 * j + 2] = 0.0f;
        dss[j] = 0.0f;
        zs[j] = 0.0f;
    }
}

// n_rays == 0: no rays, every sample slot zero, both counters zero.
__global__ void march_empty_kernel(uint32_t total_samples, uint32_t *next_loc, uint32_t *exceeded,
                                   uint32_t *__restrict__ idcs, float *__restrict__ xyzs, float *__restrict__ dirs,
                                   float *__restrict__ dss, float *__restrict__ zs) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) {
        *next_loc = 0;
        *exceeded = 0;
    }
    for (size_t j = g; j < total_samples; j += stride) {
        idcs[j] = 0u;
        xyzs[3 * j + 0] = xyzs[3 * j + 1] = xyzs[3 * j + 2] = 0.0f;
        dirs[3 * j + 0] = dirs[3 * j + 1] = dirs[3 * j + 2] = 0.0f;
        dss[j] = 0.0f;
        zs[j] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// inference
// ---------------------------------------------------------------------------

// Single workgroup: terminated slots take new ray indices in slot order
// (the reference hands them out with atomicAdd, marching.cu:300).
__global__ void __launch_bounds__(kScanThreads)
    assign_rays_kernel(uint32_t n_rays, const uint32_t *__restrict__ next_in, const bool *__restrict__ terminated,
                       const uint32_t *__restrict__ indices_in, uint32_t *__restrict__ next_out,
                       uint32_t *__restrict__ indices_out) {
    uint32_t base = *next_in;
    for (uint32_t c = 0; c < n_rays; c += kScanThreads) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t term = (i < n_rays && terminated[i]) ? 1u : 0u;
        uint32_t total;
        const uint32_t pre = block_exclusive_scan<kScanThreads>(term, &total);
        if (i < n_rays) indices_out[i] = term ? base + pre : indices_in[i];
        base += total;
    }
    if (threadIdx.x == 0) *next_out = base;
}

__global__ void __launch_bounds__(kBlock)
    march_inference_kernel(MarchParams p, bool word_ok, uint32_t n_total_rays, uint32_t n_rays, uint32_t cap,
                           const float *__restrict__ rays_o, const float *__restrict__ rays_d,
                           const float *__restrict__ t_starts, const float *__restrict__ t_ends,
                           const uint8_t *__restrict__ bits, const uint32_t *__restrict__ indices,
                           uint32_t *__restrict__ n_samples, float *__restrict__ t_out, float *__restrict__ xyzs,
                           float *__restrict__ dss, float *__restrict__ zs) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_rays) return;
    float *rx = xyzs + (size_t)i * cap * 3;
    float *rds = dss + (size_t)i * cap;
    float *rz = zs + (size_t)i * cap;
    const uint32_t r = indices[i];
    uint32_t steps = 0;
    float t = 0.0f;
    if (r < n_total_rays) {
        const float t0 = t_starts[r], t1 = t_ends[r];
        if (!(t1 < t0)) {
            const float ox = rays_o[3 * (size_t)r], oy = rays_o[3 * (size_t)r + 1], oz = rays_o[3 * (size_t)r + 2];
            const float dx = rays_d[3 * (size_t)r], dy = rays_d[3 * (size_t)r + 1], dz = rays_d[3 * (size_t)r + 2];
            const float idx_ = 1.0f / dx, idy = 1.0f / dy, idz = 1.0f / dz;
            BitCache cache = make_bit_cache(bits, word_ok);
            t = t0;
            while (steps < cap && t < t1) {
                const float px = ox + t * dx, py = oy + t * dy, pz = oz + t * dz;
                const float ds = step_size(p, t);
                float gx, gy, gz, mb;
                float nt;
                if (lookup_cell(p, bits, cache, px, py, pz, ds, gx, gy, gz, mb)) {
                    rx[3 * steps + 0] = px;
                    rx[3 * steps + 1] = py;
                    rx[3 * steps + 2] = pz;
                    rds[steps] = ds;
                    rz[steps] = t;
                    ++steps;
                    nt = t + ds;
                } else {
                    nt = skip_to(p, t, ds, cell_exit_t(p, t, px, py, pz, gx, gy, gz, mb, dx, dy, dz, idx_, idy, idz));
                }
                t = nt;
            }
            // far-plane sample (marching.cu:367-394)
            if (t >= t1) {
                const float px = ox + t1 * dx, py = oy + t1 * dy, pz = oz + t1 * dz;
                const float ds = step_size(p, t1);
                float gx, gy, gz, mb;
                if (lookup_cell(p, bits, cache, px, py, pz, ds, gx, gy, gz, mb)) {
                    if (steps > 0 && rds[steps - 1] + rz[steps - 1] >= t1) rds[steps - 1] = t1 - rz[steps - 1];
                    if (steps < cap) {
                        rx[3 * steps + 0] = px;
                        rx[3 * steps + 1] = py;
                        rx[3 * steps + 2] = pz;
                        rds[steps] = ds;
                        rz[steps] = t1;
                        ++steps;
                    } else {
                        t = t1;
                    }
                }
            }
        }
    }
    // unused sample slots are zero (the reference memsets them, marching.cu:565-570)
    for (uint32_t k = steps; k < cap; ++k) {
        rx[3 * k + 0] = 0.0f;
        rx[3 * k + 1] = 0.0f;
        rx[3 * k + 2] = 0.0f;
        rds[k] = 0.0f;
        rz[k] = 0.0f;
    }
    n_samples[i] = steps;
    t_out[i] = t;
}

// 64-bit words need an 8-byte aligned bitfield of a whole number of words.
bool bitfield_words_ok(const uint8_t *bits, uint32_t K, uint32_t G) {
    const uint64_t n_bytes = (uint64_t)K * G * G * G / 8;
    return (reinterpret_cast<uintptr_t>(bits) & 7u) == 0 && n_bytes % 8 == 0 && n_bytes > 0;
}

bool valid_march_statics(uint32_t diag, uint32_t K, uint32_t G, float bound, float portion, const char *target) {
    // Morton codes of cells in [0, G)^3 stay below G^3 only for power-of-two G
    const bool g_ok = G != 0 && G <= 1024 && (G & (G - 1)) == 0 && (uint64_t)K * G * G * G <= 0xffffffffull;
    if (diag == 0 || K == 0 || !g_ok || !(bound > 0.0f) || !(bound < 3.0e38f) || !(portion >= 0.0f) ||
        !(portion < 3.0e38f)) {
        set_error(kBadDescriptor,
                  "%s: invalid statics (diagonal_n_steps=%u K=%u G=%u bound=%g stepsize_portion=%g); need positive "
                  "diagonal_n_steps and K, G a power of two <= 1024 with K*G^3 < 2^32, finite bound > 0, finite "
                  "portion >= 0",
                  target, diag, K, G, (double)bound, (double)portion);
        return false;
    }
    return true;
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void march_rays(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "march_rays";
    MarchingDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!valid_march_statics(d.diagonal_n_steps, d.K, d.G, d.bound, d.stepsize_portion, target)) return;
    if (d.total_samples == 0) {
        set_error(kBadDescriptor, "%s: total_samples must be positive", target);
        return;
    }
    if (!check_buffer_list(buffers, {6, 7, 11, 12, 13, 14, 15}, target)) return;
    if (d.n_rays > 0 && !check_buffers(buffers, 16, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    const float *rays_o = (const float *)buffers[0];
    const float *rays_d = (const float *)buffers[1];
    const float *t_starts = (const float *)buffers[2];
    const float *t_ends = (const float *)buffers[3];
    const float *noises = (const float *)buffers[4];
    const uint8_t *bits = (const uint8_t *)buffers[5];
    uint32_t *next_loc = (uint32_t *)buffers[6];
    uint32_t *exceeded = (uint32_t *)buffers[7];
    bool *valid = (bool *)buffers[8];
    uint32_t *n_samples = (uint32_t *)buffers[9];
    uint32_t *startidx = (uint32_t *)buffers[10];
    uint32_t *idcs = (uint32_t *)buffers[11];
    float *xyzs = (float *)buffers[12];
    float *dirs = (float *)buffers[13];
    float *dss = (float *)buffers[14];
    float *zs = (float *)buffers[15];
    const MarchParams p = make_march_params(d.diagonal_n_steps, d.K, d.G, d.bound, d.stepsize_portion);
    const bool word_ok = bitfield_words_ok(bits, d.K, d.G);
    if (d.n_rays == 0) {
        hipLaunchKernelGGL(march_empty_kernel, dim3(min(div_up(d.total_samples, 256), 4096u)), dim3(256), 0, stream,
                           d.total_samples, next_loc, exceeded, idcs, xyzs, dirs, dss, zs);
        check_launch(target);
        return;
    }
    const uint32_t n_blocks = div_up(d.n_rays, kBlock);
    // count in ray-ordered chunks of 32, 256 and all remaining blocks
    const uint32_t bounds[4] = {0, std::min(n_blocks, 32u), std::min(n_blocks, 256u), n_blocks};
    for (int c = 0; c < 3; ++c) {
        if (bounds[c + 1] <= bounds[c]) continue;
        hipLaunchKernelGGL(march_count_kernel, dim3(bounds[c + 1] - bounds[c]), dim3(kBlock), 0, stream, p, word_ok,
                           d.n_rays,
                           d.total_samples, bounds[c], rays_o, rays_d, t_starts, t_ends, noises, bits, n_samples,
                           startidx);
        if (!check_launch(target)) return;
    }
    hipLaunchKernelGGL(march_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, d.n_rays, n_blocks, d.total_samples,
                       n_samples, startidx, next_loc, exceeded);
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(march_write_kernel, dim3(n_blocks), dim3(kBlock), 0, stream, p, word_ok, d.n_rays,
                       d.total_samples,
                       rays_o, rays_d, t_starts, t_ends, noises, bits, next_loc, exceeded, valid, n_samples, startidx,
                       idcs, xyzs, dirs, dss, zs);
    check_launch(target);
}

extern "C" void march_rays_inference(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "march_rays_inference";
    MarchingInferenceDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!valid_march_statics(d.diagonal_n_steps, d.K, d.G, d.bound, d.stepsize_portion, target)) return;
    if (d.march_steps_cap == 0) {
        set_error(kBadDescriptor, "%s: march_steps_cap must be positive", target);
        return;
    }
    if (d.n_rays == 0) return;
    if (!check_buffers(buffers, 15, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    const float *rays_o = (const float *)buffers[0];
    const float *rays_d = (const float *)buffers[1];
    const float *t_starts = (const float *)buffers[2];
    const float *t_ends = (const float *)buffers[3];
    const uint8_t *bits = (const uint8_t *)buffers[4];
    const uint32_t *next_in = (const uint32_t *)buffers[5];
    const bool *terminated = (const bool *)buffers[6];
    const uint32_t *indices_in = (const uint32_t *)buffers[7];
    uint32_t *next_out = (uint32_t *)buffers[8];
    uint32_t *indices_out = (uint32_t *)buffers[9];
    uint32_t *n_samples = (uint32_t *)buffers[10];
    float *t_out = (float *)buffers[11];
    float *xyzs = (float *)buffers[12];
    float *dss = (float *)buffers[13];
    float *zs = (float *)buffers[14];
    const MarchParams p = make_march_params(d.diagonal_n_steps, d.K, d.G, d.bound, d.stepsize_portion);
    hipLaunchKernelGGL(assign_rays_kernel, dim3(1), dim3(kScanThreads), 0, stream, d.n_rays, next_in, terminated,
                       indices_in, next_out, indices_out);
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(march_inference_kernel, dim3(div_up(d.n_rays, kBlock)), dim3(kBlock), 0, stream, p,
                       bitfield_words_ok(bits, d.K, d.G), d.n_total_rays, d.n_rays, d.march_steps_cap, rays_o, rays_d, t_starts, t_ends, bits,
                       indices_out, n_samples, t_out, xyzs, dss, zs);
    check_launch(target);
}
