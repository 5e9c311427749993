// integrate.hip — front-to-back volume integration with early stop.
//
// integrate_rays:           deps/volume-rendering-jax/lib/impl/integrating.cu:24-104, 325-377
// integrate_rays_backward:  integrating.cu:106-240, 379-446
// integrate_rays_inference: integrating.cu:242-322, 448-507
//
// Training integration: one wavefront per ray.  The 64 lanes load a chunk of the
// ray's samples with coalesced accesses and compute alpha = 1 - exp(-sigma ds) in
// parallel; the transmittance recurrence (T *= 1 - alpha, a running f32 product,
// so the early-stop index is exactly the oracle's) runs as a branch-free chain over
// the chunk, and the colour and depth sums run in sample order in four lanes.  Inference integration keeps one thread per pool slot
// (at most march_steps_cap samples each).  Counters are reduced per wavefront and
// added with one atomic per wave.
#include "ngp_common.h"
#include "ngp_device.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;

__global__ void zero_u32_kernel(uint32_t *p) { *p = 0u; }

__global__ void fill_zero_kernel(float *__restrict__ a, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) a[j] = 0.0f;
}

__global__ void fill_zero4_kernel(float4 *__restrict__ a, size_t n4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n4; j += stride)
        a[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Samples of a ray that lie inside the sample arrays.  Ranges produced by
// march_rays always do; a range running past total_samples is cut there instead of
// reading out of bounds.
__device__ __forceinline__ uint32_t sample_count(size_t s0, uint32_t n, uint32_t total_samples) {
    return s0 >= total_samples ? 0u : (uint32_t)min((size_t)n, (size_t)total_samples - s0);
}

__device__ __forceinline__ float4 load_drgb(const float *__restrict__ drgbs, size_t s) {
    return *reinterpret_cast<const float4 *>(drgbs + 4 * s);
}

__device__ __forceinline__ float lane_value(float v, int q) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), q));
}

// Rays of wavefront w: w, w + n_waves, w + 2 n_waves, ...  Lane i of a batch loads
// the count and start of the batch's i-th ray (one load for 64 rays), rays
// without samples are written by their own lanes at once, and the wavefront walks
// only the rays that have samples.  Training batches hold their valid rays first
// (march_rays keeps ray order), so the stride spreads them over the wavefronts.
struct RayBatch {
    uint64_t ray;  // this lane's ray
    bool in;       // ray < n_rays
    size_t s0;
    uint32_t n;    // samples inside the arrays
};

__device__ __forceinline__ RayBatch load_ray_batch(uint64_t first, uint32_t stride, uint32_t n_rays,
                                                   const uint32_t *__restrict__ startidx,
                                                   const uint32_t *__restrict__ nsamp, uint32_t total_samples) {
    RayBatch rb;
    rb.ray = first + (uint64_t)lane_id() * stride;
    rb.in = rb.ray < n_rays;
    rb.s0 = rb.in ? startidx[rb.ray] : 0;
    rb.n = rb.in ? sample_count(rb.s0, nsamp[rb.ray], total_samples) : 0u;
    return rb;
}

// A chunk of up to 64 samples of one ray, lane q holding sample c0 + q.
struct Chunk {
    float4 c;  // sigma, r, g, b
    float z, ds;
};

__device__ __forceinline__ Chunk load_chunk(const float *__restrict__ drgbs, const float *__restrict__ zs,
                                            const float *__restrict__ dss, size_t s0, uint32_t c0, uint32_t n) {
    Chunk k;
    const uint32_t j = c0 + lane_id();
    if (j < n) {
        k.c = load_drgb(drgbs, s0 + j);
        k.z = zs[s0 + j];
        k.ds = dss[s0 + j];
    } else {
        k.c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        k.z = 0.0f;
        k.ds = 0.0f;
    }
    return k;
}

// LDS rows of one wavefront: 1 - alpha, T before each sample, and the four
// product rows (r, g, b, depth).  Rows are 68 floats apart so that the four lanes
// of ordered_sums read different banks.
constexpr int kRow = kWave + 4;
constexpr int kRows = 6 * kRow;

// The transmittance recurrence over one chunk, T = T * (1 - alpha) in sample order,
// with no data-dependent branch: the 64 factors go through LDS (broadcast reads),
// lane 0 stores T before each sample, and lane q reads back T_q.  The stop index
// (first q < m with !(T_q > threshold), where the one-thread loop
// `for (; T > thr && k < n; ++k)` exits) comes from one ballot.  T returns the
// transmittance the sequential loop ends the chunk with.  Lanes >= m carry
// 1 - alpha = 1, which leaves T unchanged.
__device__ __forceinline__ int transmittance_chunk(float *rows, float om, int m, float &T, float &Tq) {
    const int lane = lane_id();
    float *om_row = rows, *t_row = rows + kRow;
    om_row[lane] = om;
    __builtin_amdgcn_wave_barrier();
    float4 om4[kWave / 4];  // all factors in registers first, so the reads overlap
#pragma unroll
    for (int q = 0; q < kWave / 4; ++q) om4[q] = reinterpret_cast<const float4 *>(om_row)[q];
    float t = T;
#pragma unroll
    for (int q = 0; q < kWave / 4; ++q) {
        const float4 o = om4[q];
        float4 v;
        v.x = t;
        t = t * o.x;
        v.y = t;
        t = t * o.y;
        v.z = t;
        t = t * o.z;
        v.w = t;
        t = t * o.w;
        if (lane == 0) reinterpret_cast<float4 *>(t_row)[q] = v;
    }
    __builtin_amdgcn_wave_barrier();
    Tq = t_row[lane];
    const uint64_t dead = __ballot(lane < m && !(Tq > kTThreshold));
    const int stop = dead ? __builtin_ctzll(dead) : m;
    T = stop < m ? lane_value(Tq, stop) : t;
    __builtin_amdgcn_wave_barrier();
    return stop;
}

// Sequential sums in sample order: lane ch (0..3) adds product row ch (64 values,
// +0 past the stop index) into acc, the running channel value, one sample at a
// time.  With `prefix`, the running value after each sample is written back into
// the row.  A sum that starts at +0 is never -0, so the +0 padding adds nothing.
__device__ __forceinline__ void ordered_sums(float *prods, float &acc, bool prefix) {
    const int lane = lane_id();
    if (lane < 4) {
        float4 *row = reinterpret_cast<float4 *>(prods + lane * kRow);
        float a = acc;
#pragma unroll 4
        for (int q = 0; q < kWave / 4; ++q) {
            float4 v = row[q];
            a = a + v.x;
            v.x = a;
            a = a + v.y;
            v.y = a;
            a = a + v.z;
            v.z = a;
            a = a + v.w;
            v.w = a;
            if (prefix) row[q] = v;
        }
        acc = a;
    }
}

// One wavefront per ray with samples.  Each chunk of 64 samples is loaded by the 64
// lanes at once (the next chunk's loads are issued before this chunk is worked);
// every lane evaluates its sample's alpha and 1 - alpha; the transmittance chain
// and the colour/depth sums then run in sample order with the one-thread-per-ray
// formulation's exact operations (w = T a, acc = acc + w c, T = T (1 - a)), so
// colours, opacities and the early-stop index equal the oracle's bit for bit.
__global__ void __launch_bounds__(kBlock)
    integrate_kernel(uint32_t n_rays, uint32_t total_samples, const uint32_t *__restrict__ startidx,
                     const uint32_t *__restrict__ nsamp, const float *__restrict__ bgs, const float *__restrict__ dss,
                     const float *__restrict__ zs, const float *__restrict__ drgbs, uint32_t *__restrict__ measured,
                     float *__restrict__ rgbd_out, float *__restrict__ opacity_out) {
    __shared__ __align__(16) float rows_all[kBlock / kWave][kRows];
    float *rows = rows_all[threadIdx.x / kWave];
    float *prods = rows + 2 * kRow;
    const int lane = lane_id();
    const uint32_t n_waves = gridDim.x * (kBlock / kWave);
    const uint32_t wave = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    uint32_t composited = 0;
    for (uint64_t first = wave; first < n_rays; first += (uint64_t)kWave * n_waves) {
        const RayBatch rb = load_ray_batch(first, n_waves, n_rays, startidx, nsamp, total_samples);
        if (rb.in && rb.n == 0) {  // T = 1: background, opacity 0
            opacity_out[rb.ray] = 0.0f;
            *reinterpret_cast<float4 *>(rgbd_out + 4 * rb.ray) =
                make_float4(0.0f + bgs[3 * rb.ray], 0.0f + bgs[3 * rb.ray + 1], 0.0f + bgs[3 * rb.ray + 2], 0.0f);
        }
        for (uint64_t todo = __ballot(rb.in && rb.n > 0); todo != 0; todo &= todo - 1) {
            const int src = __builtin_ctzll(todo);
            const uint32_t ray = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rb.ray, src);
            const size_t s0 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rb.s0, src);
            const uint32_t n = (uint32_t)__builtin_amdgcn_readlane((int)rb.n, src);
            float T = 1.0f, acc = 0.0f;  // lane ch < 4: running r, g, b, depth
            uint32_t k = 0;
            Chunk next = load_chunk(drgbs, zs, dss, s0, 0, n);
            for (uint32_t c0 = 0; c0 < n && T > kTThreshold; c0 += kWave) {
                const Chunk ch = next;
                if (c0 + kWave < n) next = load_chunk(drgbs, zs, dss, s0, c0 + kWave, n);
                const int m = (int)min(n - c0, (uint32_t)kWave);
                const float alpha = lane < m ? 1.0f - det_expf(-ch.c.x * ch.ds) : 0.0f;
                const float om = lane < m ? 1.0f - alpha : 1.0f;
                float Tq;
                const int stop = transmittance_chunk(rows, om, m, T, Tq);
                k += (uint32_t)stop;
                const float w = Tq * alpha;
                const bool in = lane < stop;
                prods[lane] = in ? w * ch.c.y : 0.0f;
                prods[kRow + lane] = in ? w * ch.c.z : 0.0f;
                prods[2 * kRow + lane] = in ? w * ch.c.w : 0.0f;
                prods[3 * kRow + lane] = in ? w * ch.z : 0.0f;
                __builtin_amdgcn_wave_barrier();
                ordered_sums(prods, acc, false);
                __builtin_amdgcn_wave_barrier();
            }
            composited += k;
            // lane ch < 4 writes channel ch
            const float op = 1.0f - T;
            if (lane == 0) opacity_out[ray] = op;
            if (lane < 4) {
                float out;
                if (T <= kTThreshold) {
                    const float inv = 1.0f / op;
                    out = acc * inv;
                } else {
                    out = lane < 3 ? acc + T * bgs[3 * (size_t)ray + lane] : acc;
                }
                rgbd_out[4 * (size_t)ray + lane] = out;
            }
        }
    }
    // one device atomic per workgroup: every wave of a batch of valid rays would
    // otherwise add to the one counter (4096 serialised atomics took ~40 us)
    __shared__ uint32_t wave_n[kBlock / kWave];
    if (lane == 0) wave_n[threadIdx.x / kWave] = composited;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < kBlock / kWave; ++w) tot += wave_n[w];
        if (tot != 0) atomicAdd(measured, tot);
    }
}

// Backward, one wavefront per ray with samples (rays batched as in the forward;
// rays without samples give only dL/dbg = gr).  The forward's chain runs again;
// lane q then holds w_q, T_{q+1} = T_q (1 - a_q) and the colour and depth sums up
// to and including sample q (the ordered prefix sums), and writes its own sample's
// gradients (integrating.cu:177-233).
__global__ void __launch_bounds__(kBlock)
    integrate_backward_kernel(uint32_t n_rays, uint32_t total_samples, float near_distance,
                              const uint32_t *__restrict__ startidx, const uint32_t *__restrict__ nsamp,
                              const float *__restrict__ bgs, const float *__restrict__ dss,
                              const float *__restrict__ zs, const float *__restrict__ drgbs,
                              const float *__restrict__ final_rgbds, const float *__restrict__ final_opacities,
                              const float *__restrict__ dL_drgbd, float *__restrict__ dL_dbgs,
                              float *__restrict__ dL_dz, float *__restrict__ dL_ddrgbs) {
    __shared__ __align__(16) float rows_all[kBlock / kWave][kRows];
    float *rows = rows_all[threadIdx.x / kWave];
    float *prods = rows + 2 * kRow;
    const int lane = lane_id();
    const uint32_t n_waves = gridDim.x * (kBlock / kWave);
    const uint32_t wave = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    for (uint64_t first = wave; first < n_rays; first += (uint64_t)kWave * n_waves) {
        const RayBatch rb = load_ray_batch(first, n_waves, n_rays, startidx, nsamp, total_samples);
        if (rb.in && rb.n == 0) {  // T = 1 > threshold: only the background gets a gradient
            const float4 gr = *reinterpret_cast<const float4 *>(dL_drgbd + 4 * rb.ray);
            dL_dbgs[3 * rb.ray + 0] = 1.0f * gr.x;
            dL_dbgs[3 * rb.ray + 1] = 1.0f * gr.y;
            dL_dbgs[3 * rb.ray + 2] = 1.0f * gr.z;
        }
        for (uint64_t todo = __ballot(rb.in && rb.n > 0); todo != 0; todo &= todo - 1) {
            const int src = __builtin_ctzll(todo);
            const uint32_t ray = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rb.ray, src);
            const size_t s0 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rb.s0, src);
            const uint32_t n = (uint32_t)__builtin_amdgcn_readlane((int)rb.n, src);
            const float4 R = *reinterpret_cast<const float4 *>(final_rgbds + 4 * (size_t)ray);
            const float4 gr = *reinterpret_cast<const float4 *>(dL_drgbd + 4 * (size_t)ray);
            const float bg0 = bgs[3 * (size_t)ray], bg1 = bgs[3 * (size_t)ray + 1], bg2 = bgs[3 * (size_t)ray + 2];
            const float op = final_opacities[ray];
            const bool terminated = op >= 1.0f - kTThreshold;
            const float bgk = terminated ? 0.0f : 1.0f - op;
            float T = 1.0f, acc = 0.0f;
            Chunk next = load_chunk(drgbs, zs, dss, s0, 0, n);
            for (uint32_t c0 = 0; c0 < n && T > kTThreshold; c0 += kWave) {
                const Chunk ch = next;
                if (c0 + kWave < n) next = load_chunk(drgbs, zs, dss, s0, c0 + kWave, n);
                const int m = (int)min(n - c0, (uint32_t)kWave);
                const float alpha = lane < m ? 1.0f - det_expf(-ch.c.x * ch.ds) : 0.0f;
                const float om = lane < m ? 1.0f - alpha : 1.0f;
                float Tq;
                const int stop = transmittance_chunk(rows, om, m, T, Tq);
                const float w = Tq * alpha;
                const float T1 = Tq * om;  // T_{i+1} enters the gradient (integrating.cu:190-192)
                const bool in = lane < stop;
                prods[lane] = in ? w * ch.c.y : 0.0f;
                prods[kRow + lane] = in ? w * ch.c.z : 0.0f;
                prods[2 * kRow + lane] = in ? w * ch.c.w : 0.0f;
                prods[3 * kRow + lane] = in ? w * ch.z : 0.0f;
                __builtin_amdgcn_wave_barrier();
                ordered_sums(prods, acc, true);
                __builtin_amdgcn_wave_barrier();
                if (in) {
                    const float my_r = prods[lane], my_g = prods[kRow + lane], my_b = prods[2 * kRow + lane],
                                my_d = prods[3 * kRow + lane];
                    const size_t s = s0 + c0 + lane;
                    dL_dz[s] = w * gr.w;
                    const float a0 = gr.x * ((T1 * ch.c.y - (R.x - my_r)) - bg0 * bgk);
                    const float a1 = gr.y * ((T1 * ch.c.z - (R.y - my_g)) - bg1 * bgk);
                    const float a2 = gr.z * ((T1 * ch.c.w - (R.z - my_b)) - bg2 * bgk);
                    const float a3 = gr.w * (T1 * ch.z - (R.w - my_d));
                    const float dsig = ch.ds * (((a0 + a1) + a2) + a3);
                    const float reg = ((double)ch.c.x > 4e-5 && ch.z < near_distance) ? 1e-4f : 0.0f;
                    const float gs = fminf(ch.z * ch.z, 1.0f);
                    *reinterpret_cast<float4 *>(dL_ddrgbs + 4 * s) =
                        make_float4(gs * dsig + reg, w * gr.x, w * gr.y, w * gr.z);
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (lane == 0) {
                const bool live = T > kTThreshold;
                dL_dbgs[3 * (size_t)ray + 0] = live ? T * gr.x : 0.0f;
                dL_dbgs[3 * (size_t)ray + 1] = live ? T * gr.y : 0.0f;
                dL_dbgs[3 * (size_t)ray + 2] = live ? T * gr.z : 0.0f;
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock)
    integrate_inference_kernel(uint32_t n_total_rays, uint32_t n_rays, uint32_t cap, const float *__restrict__ rays_bg,
                               const float *__restrict__ rays_rgbd, const float *__restrict__ rays_T,
                               const uint32_t *__restrict__ n_samples, const uint32_t *__restrict__ indices,
                               const float *__restrict__ dss, const float *__restrict__ zs,
                               const float *__restrict__ drgbs, uint32_t *__restrict__ terminate_cnt,
                               bool *__restrict__ terminated, float *__restrict__ rgbd_out, float *__restrict__ T_out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t term = 0;
    if (i < n_rays) {
        const uint32_t n = min(n_samples[i], cap);
        const uint32_t r = indices[i];
        float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float Tout = 0.0f;
        if (r < n_total_rays) {
            const size_t base = (size_t)i * cap;
            float T = rays_T[r];
            const float4 acc = *reinterpret_cast<const float4 *>(rays_rgbd + 4 * (size_t)r);
            float cr = acc.x, cg = acc.y, cb = acc.z, cd = acc.w;
            for (uint32_t k = 0; T > kTThreshold && k < n; ++k) {
                const float4 c = load_drgb(drgbs, base + k);
                const float alpha = 1.0f - det_expf(-c.x * dss[base + k]);
                const float w = T * alpha;
                cr = cr + w * c.y;
                cg = cg + w * c.z;
                cb = cb + w * c.w;
                cd = cd + w * zs[base + k];
                T = T * (1.0f - alpha);
            }
            if (T <= kTThreshold) {
                const float inv = 1.0f / (1.0f - T);
                term = 1;
                out = make_float4(cr * inv, cg * inv, cb * inv, cd * inv);
            } else {
                term = n < cap ? 1u : 0u;
                Tout = T;
                if (term) {
                    out = make_float4(cr + T * rays_bg[3 * (size_t)r], cg + T * rays_bg[3 * (size_t)r + 1],
                                      cb + T * rays_bg[3 * (size_t)r + 2], cd);
                } else {
                    out = make_float4(cr, cg, cb, cd);
                }
            }
        }
        terminated[i] = term != 0;
        *reinterpret_cast<float4 *>(rgbd_out + 4 * (size_t)i) = out;
        T_out[i] = Tout;
    }
    const uint32_t ws = wave_sum(term);
    if (lane_id() == 0 && ws != 0) atomicAdd(terminate_cnt, ws);
}

// Workgroups for one-wavefront-per-ray kernels: enough wavefronts to fill the chip
// several times over; wavefronts loop over further rays.
inline uint32_t wave_grid(uint32_t n_rays) {
    const uint32_t waves_per_block = kBlock / kWave;
    return std::max<uint32_t>(1, std::min<uint32_t>(div_up(n_rays, waves_per_block), 4096));
}

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

void zero_floats(float *p, size_t n, hipStream_t stream) {
    if (n == 0) return;
    if (aligned16(p) && (n & 3u) == 0) {
        size_t n4 = n / 4;
        uint32_t grid = (uint32_t)std::min<size_t>((n4 + 255) / 256, 8192);
        hipLaunchKernelGGL(fill_zero4_kernel, dim3(grid), dim3(256), 0, stream, reinterpret_cast<float4 *>(p), n4);
    } else {
        uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(256), 0, stream, p, n);
    }
}

// Clears two buffers with one launch when both are float4-aligned (a launch and
// its gap cost more than the few MB cleared), else one launch each.
__global__ void fill_zero4x2_kernel(float4 *__restrict__ a, size_t na4, float4 *__restrict__ b, size_t nb4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < na4 + nb4; j += stride) {
        if (j < na4) a[j] = z;
        else b[j - na4] = z;
    }
}

void zero_floats2(float *a, size_t na, float *b, size_t nb, hipStream_t stream) {
    if (aligned16(a) && aligned16(b) && (na & 3u) == 0 && (nb & 3u) == 0 && na + nb > 0) {
        const size_t n4 = (na + nb) / 4;
        const uint32_t grid = (uint32_t)std::min<size_t>((n4 + 255) / 256, 8192);
        hipLaunchKernelGGL(fill_zero4x2_kernel, dim3(grid), dim3(256), 0, stream, reinterpret_cast<float4 *>(a),
                           na / 4, reinterpret_cast<float4 *>(b), nb / 4);
    } else {
        zero_floats(a, na, stream);
        zero_floats(b, nb, stream);
    }
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void integrate_rays(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "integrate_rays";
    IntegratingDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!check_buffer_list(buffers, {6}, target)) return;
    if (d.n_rays > 0 && !check_buffers(buffers, 9, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    if (d.n_rays > 0 && (!aligned16(buffers[5]) || !aligned16(buffers[7]))) {
        set_error(kBadArgument, "%s: drgbs and final_rgbds must be 16-byte aligned", target);
        return;
    }
    uint32_t *measured = (uint32_t *)buffers[6];
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, stream, measured);
    if (d.n_rays == 0) {
        check_launch(target);
        return;
    }
    hipLaunchKernelGGL(integrate_kernel, dim3(wave_grid(d.n_rays)), dim3(kBlock), 0, stream, d.n_rays,
                       d.total_samples, (const uint32_t *)buffers[0], (const uint32_t *)buffers[1], (const float *)buffers[2],
                       (const float *)buffers[3], (const float *)buffers[4], (const float *)buffers[5], measured,
                       (float *)buffers[7], (float *)buffers[8]);
    check_launch(target);
}

extern "C" void integrate_rays_backward(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "integrate_rays_backward";
    IntegratingBackwardDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.n_rays == 0 && d.total_samples == 0) return;
    if (!check_buffer_list(buffers, d.n_rays > 0 ? std::initializer_list<int>{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11}
                                                 : std::initializer_list<int>{10, 11}, target))
        return;
    hipStream_t stream = (hipStream_t)stream_;
    if (!aligned16(buffers[11]) ||
        (d.n_rays > 0 && (!aligned16(buffers[5]) || !aligned16(buffers[6]) || !aligned16(buffers[8])))) {
        set_error(kBadArgument, "%s: drgbs, final_rgbds, dL_dfinal_rgbds and dL_ddrgbs must be 16-byte aligned",
                  target);
        return;
    }
    // samples outside every ray's composited prefix get zero gradient (integrating.cu:411-414)
    zero_floats2((float *)buffers[10], d.total_samples, (float *)buffers[11], 4 * (size_t)d.total_samples, stream);
    if (d.n_rays > 0) {
        hipLaunchKernelGGL(integrate_backward_kernel, dim3(wave_grid(d.n_rays)), dim3(kBlock), 0, stream,
                           d.n_rays, d.total_samples, d.near_distance, (const uint32_t *)buffers[0],
                           (const uint32_t *)buffers[1],
                           (const float *)buffers[2], (const float *)buffers[3], (const float *)buffers[4],
                           (const float *)buffers[5], (const float *)buffers[6], (const float *)buffers[7],
                           (const float *)buffers[8], (float *)buffers[9], (float *)buffers[10],
                           (float *)buffers[11]);
    }
    check_launch(target);
}

extern "C" void integrate_rays_inference(ngp_stream_t stream_, void **buffers, const char *opaque,
                                         size_t opaque_len) {
    const char *target = "integrate_rays_inference";
    IntegratingInferenceDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!check_buffer_list(buffers, {8}, target)) return;
    if (d.n_rays > 0 && !check_buffers(buffers, 12, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    if (d.n_rays > 0 && (!aligned16(buffers[1]) || !aligned16(buffers[7]) || !aligned16(buffers[10]))) {
        set_error(kBadArgument, "%s: rays_rgbd, drgbs and rays_rgbd_out must be 16-byte aligned", target);
        return;
    }
    uint32_t *cnt = (uint32_t *)buffers[8];
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, stream, cnt);
    if (d.n_rays > 0) {
        hipLaunchKernelGGL(integrate_inference_kernel, dim3(div_up(d.n_rays, kBlock)), dim3(kBlock), 0, stream,
                           d.n_total_rays, d.n_rays, d.march_steps_cap, (const float *)buffers[0],
                           (const float *)buffers[1], (const float *)buffers[2], (const uint32_t *)buffers[3],
                           (const uint32_t *)buffers[4], (const float *)buffers[5], (const float *)buffers[6],
                           (const float *)buffers[7], cnt, (bool *)buffers[9], (float *)buffers[10],
                           (float *)buffers[11]);
    }
    check_launch(target);
}
