// harness.hip — device-side pieces of the training step around the hot path.
//
// train_make_rays:  pixel sampling + pinhole ray generation + AABB near/far +
//                   perturbation noise + background colour + ground-truth blend
//                   in one kernel (app/nerf/train.py:47-48, app/nerf/_utils.py:93-150,
//                   utils/types.py:400-439, models/renderers/cuda.py:56-97,
//                   utils/data.py:443-464).
// train_loss_grad:  Huber(delta) loss over valid rays and its gradient with respect
//                   to final_rgbds (app/nerf/_utils.py:149-158).
// adam_update:      optax.adam(b1, b2, eps, eps_root) followed by
//                   add_decayed_weights, as chained in app/nerf/_utils.py:19-77.
//
// Random numbers come from a stateless 64-bit mix of (seed, step, ray, stream),
// so a step is reproducible from its inputs and needs no generator state.
#include <hip/hip_fp16.h>

#include "ngp_common.h"
#include "rng_device.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock)
    make_rays_kernel(NgpTrainRaysDescriptor d, const float *__restrict__ transforms, const uint8_t *__restrict__ rgba,
                     float *__restrict__ rays_o, float *__restrict__ rays_d, float *__restrict__ t_starts,
                     float *__restrict__ t_ends, float *__restrict__ noises, float *__restrict__ bgs,
                     float *__restrict__ gt, uint32_t *__restrict__ pixel_ids) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= d.n_rays) return;
    const uint64_t seed = ((uint64_t)d.seed_hi << 32) | d.seed_lo;
    const uint64_t n_px = (uint64_t)d.width * d.height;
    const uint64_t pick = draw(seed, d.step, i, 0) % (n_px * d.n_views);
    const uint32_t view = (uint32_t)(pick / n_px);
    const uint32_t pix = (uint32_t)(pick - (uint64_t)view * n_px);
    const uint32_t px = pix % d.width, py = pix / d.width;
    // camera-space direction, CV -> CG axis flip, normalised (types.py:417-439)
    const float x = (((float)px + 0.5f) - d.cx) / d.fx;
    const float y = (((float)py + 0.5f) - d.cy) / d.fy;
    const float cx = x, cy = -y, cz = -1.0f;
    const float inv = 1.0f / sqrtf(cx * cx + cy * cy + cz * cz);
    const float dcx = cx * inv, dcy = cy * inv, dcz = cz * inv;
    const float *T = transforms + 12 * (size_t)view;  // R row-major (9) then translation (3)
    const float dx = T[0] * dcx + T[1] * dcy + T[2] * dcz;
    const float dy = T[3] * dcx + T[4] * dcy + T[5] * dcz;
    const float dz = T[6] * dcx + T[7] * dcy + T[8] * dcz;
    const float ox = T[9], oy = T[10], oz = T[11];
    // AABB [-bound, bound]^3 entry / exit (cuda.py:56-97)
    const float b = d.bound;
    float t0 = 0.0f, t1 = 0.0f;
    {
        const float o3[3] = {ox, oy, oz};
        const float d3[3] = {dx, dy, dz};
        float lo = -__builtin_huge_valf(), hi = __builtin_huge_valf();
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float da = d3[a];
            da = signbit(da) ? fminf(da, -1e-15f) : fmaxf(da, 1e-15f);
            const float ta = (-b - o3[a]) / da, tb = (b - o3[a]) / da;
            lo = fmaxf(lo, fminf(ta, tb));
            hi = fminf(hi, fmaxf(ta, tb));
        }
        t0 = fmaxf(0.0f, lo);
        t1 = hi;
    }
    rays_o[3 * (size_t)i + 0] = ox;
    rays_o[3 * (size_t)i + 1] = oy;
    rays_o[3 * (size_t)i + 2] = oz;
    rays_d[3 * (size_t)i + 0] = dx;
    rays_d[3 * (size_t)i + 1] = dy;
    rays_d[3 * (size_t)i + 2] = dz;
    t_starts[i] = t0;
    t_ends[i] = t1;
    noises[i] = (d.flags & NGP_TRAIN_PERTURB) ? unit24(draw(seed, d.step, i, 1)) : 0.0f;
    float bg[3] = {d.bg[0], d.bg[1], d.bg[2]};
    if (d.flags & NGP_TRAIN_RANDOM_BG) {
        const uint64_t r = draw(seed, d.step, i, 2);
        bg[0] = (float)((uint32_t)(r >> 43) & 0x1fffffu) * (1.0f / 2097152.0f);
        bg[1] = (float)((uint32_t)(r >> 22) & 0x1fffffu) * (1.0f / 2097152.0f);
        bg[2] = (float)((uint32_t)(r >> 1) & 0x1fffffu) * (1.0f / 2097152.0f);
    }
    bgs[3 * (size_t)i + 0] = bg[0];
    bgs[3 * (size_t)i + 1] = bg[1];
    bgs[3 * (size_t)i + 2] = bg[2];
    // ground truth: rgb * alpha + bg * (1 - alpha) on [0,1] floats (data.py:443-464)
    const uint8_t *p = rgba + 4 * ((uint64_t)view * n_px + pix);
    const float alpha = (float)p[3] / 255.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) gt[3 * (size_t)i + c] = ((float)p[c] / 255.0f) * alpha + bg[c] * (1.0f - alpha);
    pixel_ids[i] = (uint32_t)(pick & 0xffffffffu);
}

// Per-ray backgrounds from a background model (app/nerf/_utils.py:128-150): the
// ground truth of each sampled pixel blended over its ray's background, and the
// background's gradient.  pixel_ids are train_make_rays' view * H * W + pixel.
__device__ __forceinline__ float pixel_alpha(const uint8_t *__restrict__ rgba, uint32_t pid, float rgb[3]) {
    const uint8_t *p = rgba + 4 * (size_t)pid;
    const float alpha = (float)p[3] / 255.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = (float)p[c] / 255.0f;
    return alpha;
}

__global__ void __launch_bounds__(kBlock)
    gt_blend_kernel(uint32_t n, const uint8_t *__restrict__ rgba, const uint32_t *__restrict__ pixel_ids,
                    const float *__restrict__ bgs, float *__restrict__ gt) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float rgb[3];
    const float alpha = pixel_alpha(rgba, pixel_ids[i], rgb);
    // the expression of make_rays_kernel (data.py:443-464)
#pragma unroll
    for (int c = 0; c < 3; ++c) gt[3 * (size_t)i + c] = rgb[c] * alpha + bgs[3 * (size_t)i + c] * (1.0f - alpha);
}

// dL/dbg = dL/dbg through the composited colour (integrate_rays_backward) plus
// dL/dgt * dgt/dbg = -dL/dpred * (1 - alpha): the Huber loss is a function of
// pred - gt, so its gradient in gt is minus its gradient in pred.
__global__ void __launch_bounds__(kBlock)
    bg_grad_kernel(uint32_t n, const uint8_t *__restrict__ rgba, const uint32_t *__restrict__ pixel_ids,
                   const float *__restrict__ dL_drgbd, const float *__restrict__ dL_dbgs, float *__restrict__ dL_dbg) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float rgb[3];
    const float one_minus_alpha = 1.0f - pixel_alpha(rgba, pixel_ids[i], rgb);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        dL_dbg[3 * (size_t)i + c] = dL_dbgs[3 * (size_t)i + c] - dL_drgbd[4 * (size_t)i + c] * one_minus_alpha;
}

__global__ void zero_u32_kernel(uint32_t *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// One workgroup counts the valid rays and writes the count and a zero loss
// accumulator, so the loss needs no separate clearing launches (the batch's flags
// are a few hundred KB; 16 bytes per load).
constexpr int kCountThreads = 1024;
__global__ void __launch_bounds__(kCountThreads) count_valid_kernel(uint32_t n, const bool *__restrict__ flags,
                                                                    uint32_t *__restrict__ count,
                                                                    float *__restrict__ loss) {
    __shared__ uint32_t wsum[kCountThreads / kWave];
    const uint8_t *f = reinterpret_cast<const uint8_t *>(flags);
    uint32_t c = 0;
    const uint32_t n16 = (reinterpret_cast<uintptr_t>(f) & 15u) == 0 ? n / 16 : 0;
    // eight 16-byte loads in flight per thread (one round for 2^17 flags)
    constexpr uint32_t kU = 8;
    for (uint32_t q0 = 0; q0 < n16; q0 += kU * kCountThreads) {
        uint4 v[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t q = q0 + u * kCountThreads + threadIdx.x;
            v[u] = q < n16 ? reinterpret_cast<const uint4 *>(f)[q] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            // bools are 0 or 1: add the bytes of each word
            const uint32_t w[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
            for (int k = 0; k < 4; ++k)
                c += (w[k] & 0xffu) + ((w[k] >> 8) & 0xffu) + ((w[k] >> 16) & 0xffu) + (w[k] >> 24);
        }
    }
    for (uint32_t i = 16 * n16 + threadIdx.x; i < n; i += kCountThreads) c += f[i] ? 1u : 0u;
    c = wave_sum(c);
    if (lane_id() == 0) wsum[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kCountThreads / kWave; ++w) t += wsum[w];
        *count = t;
        *loss = 0.0f;
    }
}

// optax.huber_loss(pred, target, delta): 0.5 e^2 if |e| <= delta else delta (|e| - delta/2)
__global__ void __launch_bounds__(kBlock)
    loss_grad_kernel(uint32_t n, float delta, const float *__restrict__ rgbd, const float *__restrict__ gt,
                     const bool *__restrict__ valid, const uint32_t *__restrict__ n_valid,
                     const NgpLossScaleState *__restrict__ scale_state, float *__restrict__ dL_drgbd,
                     float *__restrict__ loss) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float nv = (float)max(*n_valid, 1u);
        const float ls = scale_state != nullptr ? scale_state->scale : 1.0f;
        float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (valid[i]) {
            float e[3], h = 0.0f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                e[c] = rgbd[4 * (size_t)i + c] - gt[3 * (size_t)i + c];
                const float ae = fabsf(e[c]);
                h += ae <= delta ? 0.5f * e[c] * e[c] : delta * (ae - 0.5f * delta);
            }
            l = (h / 3.0f) / nv;
            const float s = 1.0f / (3.0f * nv);
            g = make_float4((fminf(fmaxf(e[0], -delta), delta) * s) * ls, (fminf(fmaxf(e[1], -delta), delta) * s) * ls,
                            (fminf(fmaxf(e[2], -delta), delta) * s) * ls, 0.0f);
        }
        *reinterpret_cast<float4 *>(dL_drgbd + 4 * (size_t)i) = g;
    }
    // wave-reduced loss
#pragma unroll
    for (int k = kWave / 2; k > 0; k >>= 1) l += __shfl_xor(l, k, kWave);
    if (lane_id() == 0 && l != 0.0f) atomicAdd(loss, l);
}

template <typename TT>
__global__ void __launch_bounds__(kBlock)
    adam_kernel(NgpAdamDescriptor d, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                float *__restrict__ v, TT *__restrict__ mirror, const NgpLossScaleState *__restrict__ scale_state) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    AdamStep a = make_adam_step(d);
    if (scale_state != nullptr) {
        if (scale_state->overflow != 0u) return;  // skipped step: nothing changes
        a.gs = a.gs / scale_state->scale;
    }
    auto step = [&](float pk, float gk_raw, float &mk, float &vk) { return adam_step(a, pk, gk_raw, mk, vk); };
    // four elements per thread and iteration with 16-byte accesses when every array
    // allows them (the table's), else one
    const bool vec = (d.n & 3u) == 0 && ((reinterpret_cast<uintptr_t>(p) | reinterpret_cast<uintptr_t>(g) |
                                          reinterpret_cast<uintptr_t>(m) | reinterpret_cast<uintptr_t>(v)) & 15u) == 0 &&
                     (mirror == nullptr || (reinterpret_cast<uintptr_t>(mirror) & (4 * sizeof(TT) - 1)) == 0);
    if (vec) {
        const size_t n4 = d.n / 4;
        for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < n4; k += stride) {
            const float4 g4 = reinterpret_cast<const float4 *>(g)[k];
            float4 p4 = reinterpret_cast<float4 *>(p)[k];
            float4 m4 = reinterpret_cast<float4 *>(m)[k];
            float4 v4 = reinterpret_cast<float4 *>(v)[k];
            p4.x = step(p4.x, g4.x, m4.x, v4.x);
            p4.y = step(p4.y, g4.y, m4.y, v4.y);
            p4.z = step(p4.z, g4.z, m4.z, v4.z);
            p4.w = step(p4.w, g4.w, m4.w, v4.w);
            reinterpret_cast<float4 *>(m)[k] = m4;
            reinterpret_cast<float4 *>(v)[k] = v4;
            reinterpret_cast<float4 *>(p)[k] = p4;
            if (mirror != nullptr) {
                mirror[4 * k + 0] = (TT)p4.x;
                mirror[4 * k + 1] = (TT)p4.y;
                mirror[4 * k + 2] = (TT)p4.z;
                mirror[4 * k + 3] = (TT)p4.w;
            }
        }
        return;
    }
    for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < d.n; k += stride) {
        float mk = m[k], vk = v[k];
        const float pk = step(p[k], g[k], mk, vk);
        m[k] = mk;
        v[k] = vk;
        p[k] = pk;
        if (mirror != nullptr) mirror[k] = (TT)pk;
    }
}

__global__ void __launch_bounds__(kBlock)
    loss_scale_check_kernel(uint64_t n, const float *__restrict__ g, NgpLossScaleState *__restrict__ state) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    bool bad = false;
    for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) bad |= !isfinite(g[k]);
    if (__ballot(bad) != 0ull && lane_id() == 0) atomicOr(&state->overflow, 1u);
}

__global__ void loss_scale_update_kernel(NgpLossScaleUpdateDescriptor d, NgpLossScaleState *__restrict__ state) {
    if (threadIdx.x != 0) return;
    NgpLossScaleState s = *state;
    if (s.overflow != 0u) {
        s.scale = fmaxf(s.scale * 0.5f, d.min_scale);
        s.good_steps = 0;
    } else if (++s.good_steps >= d.growth_interval) {
        s.scale = fminf(s.scale * 2.0f, d.max_scale);
        s.good_steps = 0;
    }
    s.overflow = 0;
    *state = s;
}

__global__ void __launch_bounds__(kBlock)
    pool_scatter_kernel(uint32_t n, uint32_t n_total, uint32_t width, const uint32_t *__restrict__ idx,
                        const float *__restrict__ src, float *__restrict__ dst) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (uint64_t)n * width) return;
    const uint32_t i = (uint32_t)(t / width), c = (uint32_t)(t % width);
    const uint32_t r = idx[i];
    if (r < n_total) dst[(size_t)r * width + c] = src[t];
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void train_make_rays(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "train_make_rays";
    NgpTrainRaysDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.n_rays == 0) return;
    if (d.n_views == 0 || d.width == 0 || d.height == 0 || !(d.bound > 0.0f)) {
        set_error(kBadDescriptor, "%s: need views, a positive image size and bound", target);
        return;
    }
    if (!check_buffers(buffers, 10, target)) return;
    hipLaunchKernelGGL(make_rays_kernel, dim3(div_up(d.n_rays, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, d,
                       (const float *)buffers[0], (const uint8_t *)buffers[1], (float *)buffers[2],
                       (float *)buffers[3], (float *)buffers[4], (float *)buffers[5], (float *)buffers[6],
                       (float *)buffers[7], (float *)buffers[8], (uint32_t *)buffers[9]);
    check_launch(target);
}

extern "C" void train_loss_grad(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "train_loss_grad";
    NgpLossDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.flags & ~NGP_LOSS_SCALED) {
        set_error(kBadDescriptor, "%s: unknown flags %u", target, d.flags);
        return;
    }
    const bool scaled = (d.flags & NGP_LOSS_SCALED) != 0;
    if (!check_buffer_list(buffers, {3, 5}, target)) return;
    if (d.n_rays > 0 && !check_buffers(buffers, 6, target)) return;
    if (scaled && !check_buffer_list(buffers, {6}, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    uint32_t *n_valid = (uint32_t *)buffers[3];
    float *loss = (float *)buffers[5];
    hipLaunchKernelGGL(count_valid_kernel, dim3(1), dim3(kCountThreads), 0, stream, d.n_rays,
                       (const bool *)buffers[2], n_valid, loss);
    if (d.n_rays > 0) {
        hipLaunchKernelGGL(loss_grad_kernel, dim3(div_up(d.n_rays, kBlock)), dim3(kBlock), 0, stream, d.n_rays,
                           d.delta, (const float *)buffers[0], (const float *)buffers[1], (const bool *)buffers[2],
                           n_valid, scaled ? (const NgpLossScaleState *)buffers[6] : nullptr, (float *)buffers[4],
                           loss);
    }
    check_launch(target);
}

extern "C" void train_gt_blend(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "train_gt_blend";
    NgpTrainBgDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.n_rays == 0) return;
    if (!check_buffers(buffers, 4, target)) return;
    hipLaunchKernelGGL(gt_blend_kernel, dim3(div_up(d.n_rays, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, d.n_rays,
                       (const uint8_t *)buffers[0], (const uint32_t *)buffers[1], (const float *)buffers[2],
                       (float *)buffers[3]);
    check_launch(target);
}

extern "C" void train_bg_grad(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "train_bg_grad";
    NgpTrainBgDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.n_rays == 0) return;
    if (!check_buffers(buffers, 5, target)) return;
    hipLaunchKernelGGL(bg_grad_kernel, dim3(div_up(d.n_rays, kBlock)), dim3(kBlock), 0, (hipStream_t)stream, d.n_rays,
                       (const uint8_t *)buffers[0], (const uint32_t *)buffers[1], (const float *)buffers[2],
                       (const float *)buffers[3], (float *)buffers[4]);
    check_launch(target);
}

extern "C" void adam_update(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "adam_update";
    NgpAdamDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.flags & ~NGP_LOSS_SCALED) {
        set_error(kBadDescriptor, "%s: unknown flags %u", target, d.flags);
        return;
    }
    if (d.n == 0) return;
    if (!check_buffers(buffers, 4, target)) return;
    const bool scaled = (d.flags & NGP_LOSS_SCALED) != 0;
    if (scaled && !check_buffer_list(buffers, {5}, target)) return;
    const NgpLossScaleState *state = scaled ? (const NgpLossScaleState *)buffers[5] : nullptr;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(d.n, kBlock), 8192);
    if (d.mirror_dtype == NGP_DTYPE_F16 && buffers[4] != nullptr) {
        hipLaunchKernelGGL(adam_kernel<__half>, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, d,
                           (float *)buffers[0], (const float *)buffers[1], (float *)buffers[2], (float *)buffers[3],
                           (__half *)buffers[4], state);
    } else {
        hipLaunchKernelGGL(adam_kernel<float>, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, d,
                           (float *)buffers[0], (const float *)buffers[1], (float *)buffers[2], (float *)buffers[3],
                           (float *)nullptr, state);
    }
    check_launch(target);
}

extern "C" void loss_scale_check(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "loss_scale_check";
    NgpLossScaleCheckDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!check_buffer_list(buffers, {1}, target)) return;
    if (d.n == 0) return;
    if (!check_buffers(buffers, 2, target)) return;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(div_up(d.n, kBlock), 4096);
    hipLaunchKernelGGL(loss_scale_check_kernel, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, d.n,
                       (const float *)buffers[0], (NgpLossScaleState *)buffers[1]);
    check_launch(target);
}

extern "C" void loss_scale_update(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "loss_scale_update";
    NgpLossScaleUpdateDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.growth_interval == 0 || !(d.min_scale > 0.0f) || !(d.max_scale >= d.min_scale)) {
        set_error(kBadDescriptor, "%s: need growth_interval > 0 and 0 < min_scale <= max_scale", target);
        return;
    }
    if (!check_buffers(buffers, 1, target)) return;
    hipLaunchKernelGGL(loss_scale_update_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, d,
                       (NgpLossScaleState *)buffers[0]);
    check_launch(target);
}

extern "C" void pool_scatter(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "pool_scatter";
    NgpPoolScatterDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.width == 0) {
        set_error(kBadDescriptor, "%s: width must be positive", target);
        return;
    }
    if (d.n == 0) return;
    if (!check_buffers(buffers, 3, target)) return;
    const uint64_t total = (uint64_t)d.n * d.width;
    hipLaunchKernelGGL(pool_scatter_kernel, dim3((uint32_t)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, d.n, d.n_total, d.width, (const uint32_t *)buffers[0],
                       (const float *)buffers[1], (float *)buffers[2]);
    check_launch(target);
}
