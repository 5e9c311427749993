// integrate.hip — front-to-back volume integration with early stop.
//
// integrate_rays:           deps/volume-rendering-jax/lib/impl/integrating.cu:24-104, 325-377
// integrate_rays_backward:  integrating.cu:106-240, 379-446
// integrate_rays_inference: integrating.cu:242-322, 448-507
//
// Training integration: one wavefront per ray.  The 64 lanes load a chunk of the
// ray's samples with coalesced accesses and compute alpha = 1 - exp(-sigma ds) in
// parallel; the serial transmittance recurrence (T *= 1 - alpha, a running f32
// product, so the early-stop index is exactly the oracle's) then walks the chunk on
// wavefront-uniform values.  Inference integration keeps one thread per pool slot
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

// One wavefront per ray (rays assigned round-robin to the grid's wavefronts).
// Each chunk of 64 samples is loaded by the 64 lanes at once and every lane
// evaluates its sample's alpha; the compositing recurrence then walks the chunk in
// order on wavefront-uniform values, with exactly the operation order of the
// one-thread-per-ray formulation (so the early-stop index is the oracle's).
__global__ void __launch_bounds__(kBlock)
    integrate_kernel(uint32_t n_rays, uint32_t total_samples, const uint32_t *__restrict__ startidx,
                     const uint32_t *__restrict__ nsamp, const float *__restrict__ bgs, const float *__restrict__ dss,
                     const float *__restrict__ zs, const float *__restrict__ drgbs, uint32_t *__restrict__ measured,
                     float *__restrict__ rgbd_out, float *__restrict__ opacity_out) {
    const int lane = lane_id();
    const uint32_t n_waves = gridDim.x * (kBlock / kWave);
    uint32_t composited = 0;
    for (uint32_t ray = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; ray < n_rays; ray += n_waves) {
        const size_t s0 = startidx[ray];
        const uint32_t n = sample_count(s0, nsamp[ray], total_samples);
        float T = 1.0f, r = 0.0f, g = 0.0f, b = 0.0f, depth = 0.0f;
        uint32_t k = 0;
        for (uint32_t c0 = 0; c0 < n && T > kTThreshold; c0 += kWave) {
            const uint32_t j = c0 + lane;
            float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float z = 0.0f, alpha = 0.0f;
            if (j < n) {
                c = load_drgb(drgbs, s0 + j);
                z = zs[s0 + j];
                alpha = 1.0f - det_expf(-c.x * dss[s0 + j]);
            }
            const int m = (int)min(n - c0, (uint32_t)kWave);
            for (int q = 0; q < m && T > kTThreshold; ++q) {
                const float a = lane_value(alpha, q);
                const float w = T * a;
                r = r + w * lane_value(c.y, q);
                g = g + w * lane_value(c.z, q);
                b = b + w * lane_value(c.w, q);
                depth = depth + w * lane_value(z, q);
                T = T * (1.0f - a);
                ++k;
            }
        }
        composited += k;
        if (lane == 0) {
            const float op = 1.0f - T;
            opacity_out[ray] = op;
            float4 out;
            if (T <= kTThreshold) {
                const float inv = 1.0f / op;
                out = make_float4(r * inv, g * inv, b * inv, depth * inv);
            } else {
                out = make_float4(r + T * bgs[3 * (size_t)ray], g + T * bgs[3 * (size_t)ray + 1],
                                  b + T * bgs[3 * (size_t)ray + 2], depth);
            }
            *reinterpret_cast<float4 *>(rgbd_out + 4 * (size_t)ray) = out;
        }
    }
    if (lane == 0 && composited != 0) atomicAdd(measured, composited);
}

// Backward, one wavefront per ray.  The recurrence runs as in the forward; lane q
// keeps the state right after sample q (w_q, T_{q+1}, running colour and depth),
// then every lane writes its own sample's gradients (integrating.cu:177-233).
__global__ void __launch_bounds__(kBlock)
    integrate_backward_kernel(uint32_t n_rays, uint32_t total_samples, float near_distance,
                              const uint32_t *__restrict__ startidx, const uint32_t *__restrict__ nsamp,
                              const float *__restrict__ bgs, const float *__restrict__ dss,
                              const float *__restrict__ zs, const float *__restrict__ drgbs,
                              const float *__restrict__ final_rgbds, const float *__restrict__ final_opacities,
                              const float *__restrict__ dL_drgbd, float *__restrict__ dL_dbgs,
                              float *__restrict__ dL_dz, float *__restrict__ dL_ddrgbs) {
    const int lane = lane_id();
    const uint32_t n_waves = gridDim.x * (kBlock / kWave);
    for (uint32_t ray = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; ray < n_rays; ray += n_waves) {
        const size_t s0 = startidx[ray];
        const uint32_t n = sample_count(s0, nsamp[ray], total_samples);
        const float4 R = *reinterpret_cast<const float4 *>(final_rgbds + 4 * (size_t)ray);
        const float4 gr = *reinterpret_cast<const float4 *>(dL_drgbd + 4 * (size_t)ray);
        const float bg0 = bgs[3 * (size_t)ray], bg1 = bgs[3 * (size_t)ray + 1], bg2 = bgs[3 * (size_t)ray + 2];
        const float op = final_opacities[ray];
        const bool terminated = op >= 1.0f - kTThreshold;
        const float bgk = terminated ? 0.0f : 1.0f - op;
        float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f, cd = 0.0f;
        for (uint32_t c0 = 0; c0 < n && T > kTThreshold; c0 += kWave) {
            const uint32_t j = c0 + lane;
            float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float z = 0.0f, ds = 0.0f, alpha = 0.0f;
            if (j < n) {
                c = load_drgb(drgbs, s0 + j);
                z = zs[s0 + j];
                ds = dss[s0 + j];
                alpha = 1.0f - det_expf(-c.x * ds);
            }
            const int m = (int)min(n - c0, (uint32_t)kWave);
            float my_w = 0.0f, my_T = 0.0f, my_r = 0.0f, my_g = 0.0f, my_b = 0.0f, my_d = 0.0f;
            int done = 0;  // samples of this chunk composited
            for (int q = 0; q < m && T > kTThreshold; ++q) {
                const float a = lane_value(alpha, q);
                const float w = T * a;
                cr = cr + w * lane_value(c.y, q);
                cg = cg + w * lane_value(c.z, q);
                cb = cb + w * lane_value(c.w, q);
                cd = cd + w * lane_value(z, q);
                T = T * (1.0f - a);  // T_{i+1} enters the gradient (integrating.cu:190-192)
                if (lane == q) {
                    my_w = w;
                    my_T = T;
                    my_r = cr;
                    my_g = cg;
                    my_b = cb;
                    my_d = cd;
                }
                done = q + 1;
            }
            if (lane < done) {
                const size_t s = s0 + j;
                dL_dz[s] = my_w * gr.w;
                const float a0 = gr.x * ((my_T * c.y - (R.x - my_r)) - bg0 * bgk);
                const float a1 = gr.y * ((my_T * c.z - (R.y - my_g)) - bg1 * bgk);
                const float a2 = gr.z * ((my_T * c.w - (R.z - my_b)) - bg2 * bgk);
                const float a3 = gr.w * (my_T * z - (R.w - my_d));
                const float dsig = ds * (((a0 + a1) + a2) + a3);
                const float reg = ((double)c.x > 4e-5 && z < near_distance) ? 1e-4f : 0.0f;
                const float gs = fminf(z * z, 1.0f);
                *reinterpret_cast<float4 *>(dL_ddrgbs + 4 * s) =
                    make_float4(gs * dsig + reg, my_w * gr.x, my_w * gr.y, my_w * gr.z);
            }
        }
        if (lane == 0) {
            const bool live = T > kTThreshold;
            dL_dbgs[3 * (size_t)ray + 0] = live ? T * gr.x : 0.0f;
            dL_dbgs[3 * (size_t)ray + 1] = live ? T * gr.y : 0.0f;
            dL_dbgs[3 * (size_t)ray + 2] = live ? T * gr.z : 0.0f;
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
    zero_floats((float *)buffers[10], d.total_samples, stream);
    zero_floats((float *)buffers[11], 4 * (size_t)d.total_samples, stream);
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
