// proposal.hip — the learned proposal network (no counterpart in the reference):
// a small hash grid (L = 5) and a 10 -> 16 -> 1 MLP predict density at 64 equal
// bins of each ray's segment [t0, t1); trained with a histogram-bound loss against
// the NeRF's own sample weights w_i = T_i alpha_i (integrating.cu:58-66), it tells
// where each ray's mass lies.
//
//   proposal_points:       bin centres t0 + (b + 1/2) D, D = (t1 - t0) / 64
//   proposal_mlp_forward:  sigma = exp(relu(x W1) W2) per point, f32 math, weights
//                          f32 or f16 (the MLP has 176 weights and K = 10: plain
//                          FMAs in registers, MFMA tiles would be mostly padding)
//   proposal_mlp_backward: dL/dx and dL/dW (fixed-order reductions)
//   proposal_loss:         per ray, one wavefront = 64 lanes = 64 bins
#include <hip/hip_fp16.h>

#include "ngp_common.h"
#include "ngp_device.h"
#include "hashgrid_device.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;
constexpr int kBins = NGP_PROPOSAL_BINS;
constexpr int kIn = 10, kHid = 16;
constexpr int kParams = NGP_PROPOSAL_N_PARAMS;  // W1 [10][16], W2 [16][1]
static_assert(kBins == kWave, "one lane per bin");
static_assert(kParams == kIn * kHid + kHid, "parameter count");

__global__ void __launch_bounds__(kBlock)
    proposal_points_kernel(uint32_t n_rays, const float *__restrict__ rays_o, const float *__restrict__ rays_d,
                           const float *__restrict__ t0s, const float *__restrict__ t1s, float *__restrict__ pts) {
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (uint64_t)n_rays * kBins) return;
    const uint32_t r = (uint32_t)(t / kBins), b = (uint32_t)(t % kBins);
    const float t0 = t0s[r], t1 = t1s[r];
    const float D = (t1 - t0) / (float)kBins;
    const float tc = t0 + ((float)b + 0.5f) * D;
#pragma unroll
    for (int c = 0; c < 3; ++c) pts[3 * t + c] = rays_o[3 * (size_t)r + c] + tc * rays_d[3 * (size_t)r + c];
}

template <typename PT>
__device__ __forceinline__ void stage_params(float *w, const PT *__restrict__ params) {
    for (int k = threadIdx.x; k < kParams; k += kBlock) w[k] = (float)params[k];
    __syncthreads();
}

__device__ __forceinline__ void load_x(const float *__restrict__ enc, bool level_major, uint32_t n, uint32_t i,
                                       float x[kIn]) {
#pragma unroll
    for (int l = 0; l < kIn / 2; ++l) {
        const size_t at = level_major ? ((size_t)l * n + i) * 2 : (size_t)i * kIn + 2 * l;
        x[2 * l] = enc[at];
        x[2 * l + 1] = enc[at + 1];
    }
}

__device__ __forceinline__ float hidden_and_logit(const float *w, const float x[kIn], float h[kHid]) {
    float o = 0.0f;
#pragma unroll
    for (int j = 0; j < kHid; ++j) {
        float a = 0.0f;
#pragma unroll
        for (int k = 0; k < kIn; ++k) a = a + x[k] * w[k * kHid + j];
        h[j] = fmaxf(a, 0.0f);
        o = o + h[j] * w[kIn * kHid + j];
    }
    return o;
}

template <typename PT>
__global__ void __launch_bounds__(kBlock)
    proposal_fwd_kernel(uint32_t n, bool level_major, const float *__restrict__ enc, const PT *__restrict__ params,
                        float *__restrict__ sigma) {
    __shared__ float w[kParams];
    stage_params(w, params);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        float x[kIn], h[kHid];
        load_x(enc, level_major, n, i, x);
        sigma[i] = det_expf(hidden_and_logit(w, x, h));
    }
}

// One workspace row of kParams per workgroup; waves add in a fixed order.
// One halving step over N values per lane: lanes with bit d set keep values
// [N/2, N), the others [0, N/2), each adding its partner's (lane ^ d) copy of them.
template <int N>
__device__ __forceinline__ void halve_lanes(float *v, int lane, int d) {
    const bool hi = (lane & d) != 0;
#pragma unroll
    for (int k = 0; k < N / 2; ++k) {
        const float keep = hi ? v[k + N / 2] : v[k];
        const float give = hi ? v[k] : v[k + N / 2];
        v[k] = keep + __shfl_xor(give, d, kWave);
    }
}

template <typename PT>
__global__ void __launch_bounds__(kBlock)
    proposal_bwd_kernel(uint32_t n, bool level_major, const float *__restrict__ enc, const PT *__restrict__ params,
                        const float *__restrict__ dL_dsigma, float *__restrict__ dL_denc,
                        float *__restrict__ workspace) {
    __shared__ float w[kParams];
    __shared__ float part[kBlock / kWave][kParams];
    stage_params(w, params);
    float gW[kParams];
#pragma unroll
    for (int k = 0; k < kParams; ++k) gW[k] = 0.0f;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        float x[kIn], h[kHid];
        load_x(enc, level_major, n, i, x);
        const float o = hidden_and_logit(w, x, h);
        const float dz = dL_dsigma[i] * det_expf(fminf(fmaxf(o, -15.0f), 15.0f));
        float dx[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) dx[k] = 0.0f;
#pragma unroll
        for (int j = 0; j < kHid; ++j) {
            gW[kIn * kHid + j] += h[j] * dz;
            const float dh = h[j] > 0.0f ? dz * w[kIn * kHid + j] : 0.0f;
#pragma unroll
            for (int k = 0; k < kIn; ++k) {
                gW[k * kHid + j] += x[k] * dh;
                dx[k] = dx[k] + dh * w[k * kHid + j];
            }
        }
#pragma unroll
        for (int l = 0; l < kIn / 2; ++l) {
            const size_t at = level_major ? ((size_t)l * n + i) * 2 : (size_t)i * kIn + 2 * l;
            dL_denc[at] = dx[2 * l];
            dL_denc[at + 1] = dx[2 * l + 1];
        }
    }
    // the wavefront's sums of the 176 gradients by recursive halving: at distance d
    // (32, 16, 8, 4) a lane keeps the half of its remaining parameters that its lane
    // bit d picks and adds its partner's partial sums of that half (one exchange per
    // kept value instead of six per parameter); the 11 left per lane are summed over
    // the lanes that differ in bits 0-1, and lane 4m holds parameters [off, off + 11)
    static_assert(kParams == 176, "halving 176 -> 11");
    const int wave = (int)threadIdx.x / kWave, lane = lane_id();
    halve_lanes<176>(gW, lane, 32);
    halve_lanes<88>(gW, lane, 16);
    halve_lanes<44>(gW, lane, 8);
    halve_lanes<22>(gW, lane, 4);
#pragma unroll
    for (int k = 0; k < 11; ++k) {
        gW[k] += __shfl_xor(gW[k], 2, kWave);
        gW[k] += __shfl_xor(gW[k], 1, kWave);
    }
    if ((lane & 3) == 0) {
        const int off = ((lane >> 5) & 1) * 88 + ((lane >> 4) & 1) * 44 + ((lane >> 3) & 1) * 22 + ((lane >> 2) & 1) * 11;
#pragma unroll
        for (int k = 0; k < 11; ++k) part[wave][off + k] = gW[k];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kParams; k += kBlock) {
        float s = 0.0f;
        for (int v = 0; v < kBlock / kWave; ++v) s += part[v][k];
        workspace[(size_t)blockIdx.x * kParams + k] = s;
    }
}

// Column k of the workspace summed over its rows: one workgroup per column, a
// strided partial sum per thread, the wavefront sums, then the wavefronts' in order
// (a thread per column walking all rows waited on one load per row: 58 us at 512 rows).
__global__ void __launch_bounds__(kBlock)
    proposal_sum_rows_kernel(uint32_t rows, const float *__restrict__ ws, float *__restrict__ out) {
    __shared__ float part[kBlock / kWave];
    const uint32_t k = blockIdx.x;
    float s = 0.0f;
    for (uint32_t r = threadIdx.x; r < rows; r += kBlock) s += ws[(size_t)r * kParams + k];
    s = wave_sumf(s);
    if (lane_id() == 0) part[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int w = 0; w < kBlock / kWave; ++w) t += part[w];
        out[k] = t;
    }
}

// Inclusive product scan over the wavefront.
__device__ __forceinline__ float wave_prod_scan(float v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float o = __shfl_up(v, d, kWave);
        if (lane >= d) v *= o;
    }
    return v;
}

// Inclusive suffix sum over the wavefront (lane b gets sum over b' >= b).
__device__ __forceinline__ float wave_suffix_sum(float v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float o = __shfl_down(v, d, kWave);
        if (lane + d < kWave) v += o;
    }
    return v;
}

constexpr float kTThreshold = 1e-4f;  // integrating.cu:12

__global__ void __launch_bounds__(kBlock)
    proposal_loss_kernel(uint32_t n_rays, float eps, const uint32_t *__restrict__ startidx,
                         const uint32_t *__restrict__ nsamp, const float *__restrict__ dss,
                         const float *__restrict__ z_vals, const float *__restrict__ drgbs,
                         const float *__restrict__ t0s, const float *__restrict__ t1s,
                         const float *__restrict__ sigma_p, const bool *__restrict__ valid,
                         const uint32_t *__restrict__ n_valid, float *__restrict__ dL_dsigma_p,
                         float *__restrict__ loss) {
    __shared__ float bins[kBlock / kWave][kBins];
    const int wave = (int)threadIdx.x / kWave, lane = lane_id();
    const uint32_t r = blockIdx.x * (kBlock / kWave) + wave;
    if (r >= n_rays) return;
    float *out = dL_dsigma_p + (size_t)r * kBins;
    if (!valid[r]) {
        out[lane] = 0.0f;
        return;
    }
    const float nv = (float)max(*n_valid, 1u);
    const float t0 = t0s[r], t1 = t1s[r];
    const float inv = (float)kBins / (t1 - t0);
    const float D = (t1 - t0) / (float)kBins;
    bins[wave][lane] = 0.0f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t s0 = startidx[r], m = nsamp[r];
    float T = 1.0f;
    for (uint32_t c = 0; c < m; c += kWave) {
        const uint32_t k = c + lane;
        const bool live = k < m;
        const float a = live ? 1.0f - det_expf(-drgbs[4 * (size_t)(s0 + k)] * dss[s0 + k]) : 0.0f;
        const float incl = wave_prod_scan(1.0f - a);
        float excl = __shfl_up(incl, 1, kWave);
        if (lane == 0) excl = 1.0f;
        const float Tk = T * excl;
        if (live && Tk > kTThreshold) {
            const float x = (z_vals[s0 + k] - t0) * inv;
            const int b = (int)fminf(fmaxf(floorf(x), 0.0f), (float)(kBins - 1));
            atomicAdd(&bins[wave][b], Tk * a);
        }
        T = T * __shfl(incl, kWave - 1, kWave);
        if (T <= kTThreshold) break;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float W = bins[wave][lane];
    const float a = 1.0f - det_expf(-sigma_p[(size_t)r * kBins + lane] * D);
    const float incl = wave_prod_scan(1.0f - a);
    float Tb = __shfl_up(incl, 1, kWave);
    if (lane == 0) Tb = 1.0f;
    const float wp = Tb * a;
    const float gap = fmaxf(W - wp, 0.0f);
    const float l = gap * gap / (W + eps) / nv;
    const float g = -2.0f * gap / (W + eps) / nv;
    const float suffix = wave_suffix_sum(g * wp) - g * wp;
    out[lane] = D * (g * (Tb * (1.0f - a)) - suffix);
    const float ls = wave_sumf(l);
    if (lane == 0 && ls != 0.0f) atomicAdd(loss, ls);
}

// Inclusive prefix sum over the wavefront.
__device__ __forceinline__ float wave_prefix_sum(float v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float o = __shfl_up(v, d, kWave);
        if (lane >= d) v += o;
    }
    return v;
}

struct SampleArgs {
    uint32_t n_per_ray, mode;
    float mass_threshold, eps;
    MarchParams mp;
    uint32_t n_uniform = 0;  // training: stratified uniform samples added to each ray
    float empty_mass = 0.0f;  // below it a ray renders as background, without fallback
};

// Ascending bitonic sort of one value per lane across the wavefront.
__device__ __forceinline__ float wave_sort_ascending(float v) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const float o = __shfl_xor(v, j, kWave);
            const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
            v = keep_min ? fminf(v, o) : fmaxf(v, o);
        }
    }
    return v;
}

// Ray r's sample plan from lane b's bin density sigma_b (one wavefront per ray,
// lane b = bin b; lane j < n_per_ray plans sample j): its t and ds, its position,
// and whether the occupancy grid marks its cell occupied (march_rays never samples
// empty cells).  Quantile placement puts sample j at the (j + jitter) / n_per_ray
// quantile of the piecewise constant density w_b / (M D): jitter 1/2 renders,
// a per-ray uniform jitter stratifies training samples.
struct SamplePlan {
    float t, ds, p[3];
    bool keep;  // this lane's sample exists and its cell is occupied
    bool ok;     // the ray has a segment and enough predicted mass
    bool seg;    // t1 > t0
    bool empty;  // predicted mass below empty_mass: background, no fallback
};

__device__ __forceinline__ SamplePlan plan_samples(const SampleArgs &sa, uint32_t r, float sigma_b, float t0, float t1,
                                                   float jitter, const float *__restrict__ rays_o,
                                                   const float *__restrict__ rays_d, const uint8_t *__restrict__ bits) {
    const int lane = lane_id();
    const uint32_t N = sa.n_per_ray;
    SamplePlan pl;
    pl.seg = t1 > t0;
    const float D = (t1 - t0) / (float)kBins;
    const float a = pl.seg ? 1.0f - det_expf(-sigma_b * D) : 0.0f;
    const float incl = wave_prod_scan(1.0f - a);
    float Tb = __shfl_up(incl, 1, kWave);
    if (lane == 0) Tb = 1.0f;
    const float w = Tb * a;
    const float M = wave_sumf(w);
    pl.ok = pl.seg && M >= sa.mass_threshold && M > 0.0f;
    pl.empty = pl.seg && !pl.ok && M < sa.empty_mass;
    float t = 0.0f, ds = 0.0f;
    if (pl.ok && sa.mode == NGP_PROPOSAL_SUPPORT) {
        // bins [lo, hi]: first with cumulative share > eps, first with >= 1 - eps
        const float inc = wave_prefix_sum(w / M);
        const uint64_t above = __ballot(inc > sa.eps), full = __ballot(inc >= 1.0f - sa.eps);
        const int lo = above ? __builtin_ctzll(above) : 0;
        int hi = full ? __builtin_ctzll(full) : kBins - 1;
        hi = hi > lo ? hi : lo;
        const float t_lo = t0 + (float)lo * D, t_hi = t0 + (float)(hi + 1) * D;
        ds = (t_hi - t_lo) / (float)N;
        t = t_lo + ((float)lane + 0.5f) * ds;
    } else if (pl.ok) {
        const float p = w / M;
        const float C = wave_prefix_sum(p) - p;  // exclusive
        const float u = ((float)lane + jitter) / (float)N;
        // last bin with C_b <= u (C is non-decreasing, C_0 = 0)
        int b = 0;
#pragma unroll
        for (int step = kBins / 2; step > 0; step >>= 1) {
            const int cand = b + step;
            const float cc = __shfl(C, cand < kBins ? cand : kBins - 1, kWave);
            if (cand < kBins && cc <= u) b = cand;
        }
        const float pb = __shfl(p, b, kWave), cb = __shfl(C, b, kWave);
        const float frac = pb > 0.0f ? fminf(fmaxf((u - cb) / pb, 0.0f), 1.0f) : 0.5f;
        t = t0 + ((float)b + frac) * D;
        const float tn = __shfl_down(t, 1, kWave);
        ds = (uint32_t)lane + 1 < N ? tn - t : fminf(pb > 0.0f ? D / ((float)N * pb) : D, t1 - t);
    }
    // Training: U more samples stratified uniformly over [t0, t1) (lane N + i at
    // t0 + (i + jitter) (t1 - t0) / U), so the field also sees where the proposal
    // predicts no mass; the N + U samples are sorted and each sample's ds is the
    // distance to the next (the last: up to t1, at most the uniform spacing).
    const uint32_t U = sa.n_uniform;
    bool live = pl.ok && (uint32_t)lane < N;
    if (U > 0) {
        const float span = (t1 - t0) / (float)U;
        const bool uni = pl.seg && (uint32_t)lane >= N && (uint32_t)lane < N + U;
        const float key = live ? t : (uni ? t0 + ((float)(lane - (int)N) + jitter) * span : __builtin_huge_valf());
        t = wave_sort_ascending(key);
        live = t < __builtin_huge_valf();
        const float tn = __shfl_down(t, 1, kWave);
        const bool last = (uint32_t)lane + 1 == kWave || !(tn < __builtin_huge_valf());
        ds = live ? (last ? fminf(t1 - t, span) : tn - t) : 0.0f;
        if (!live) t = 0.0f;
        pl.ok = pl.seg;  // the ray has samples whatever its predicted mass
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) pl.p[c] = live ? rays_o[3 * (size_t)r + c] + t * rays_d[3 * (size_t)r + c] : 0.0f;
    pl.keep = false;
    if (live) {
        BitCache cache = make_bit_cache(bits, false);
        float gx, gy, gz, mb;
        pl.keep = lookup_cell(sa.mp, bits, cache, pl.p[0], pl.p[1], pl.p[2], step_size(sa.mp, t), gx, gy, gz, mb);
    }
    pl.t = t;
    pl.ds = ds;
    return pl;
}

// Places ray r's samples: the kept ones packed in order at the front of the ray's
// slots, counts[r] of them, the remaining slots zero; a ray below the mass
// threshold gets none and fallback = true.
__device__ __forceinline__ void place_samples(const SampleArgs &sa, uint32_t r, float sigma_b, float t0, float t1,
                                              const float *__restrict__ rays_o, const float *__restrict__ rays_d,
                                              const uint8_t *__restrict__ bits, uint32_t *__restrict__ startidx,
                                              uint32_t *__restrict__ counts, bool *__restrict__ fallback,
                                              float *__restrict__ xyzs, float *__restrict__ dss,
                                              float *__restrict__ zs) {
    const int lane = lane_id();
    const uint32_t N = sa.n_per_ray;
    const SamplePlan pl = plan_samples(sa, r, sigma_b, t0, t1, 0.5f, rays_o, rays_d, bits);
    const uint64_t kept = __ballot(pl.keep);
    const uint32_t n_kept = (uint32_t)__popcll(kept);
    const size_t base = (size_t)r * N;
    if (lane == 0) {
        startidx[r] = (uint32_t)base;
        counts[r] = n_kept;
        fallback[r] = pl.seg && !pl.ok && !pl.empty;
    }
    if (pl.keep) {
        const size_t j = base + (uint32_t)__popcll(kept & ((1ull << lane) - 1ull));
#pragma unroll
        for (int c = 0; c < 3; ++c) xyzs[3 * j + c] = pl.p[c];
        dss[j] = pl.ds;
        zs[j] = pl.t;
    }
    if ((uint32_t)lane >= n_kept && (uint32_t)lane < N) {
        const size_t j = base + lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) xyzs[3 * j + c] = 0.0f;
        dss[j] = 0.0f;
        zs[j] = 0.0f;
    }
}

__global__ void __launch_bounds__(kBlock)
    proposal_sample_kernel(uint32_t n_rays, SampleArgs sa, const float *__restrict__ rays_o,
                           const float *__restrict__ rays_d, const float *__restrict__ t0s,
                           const float *__restrict__ t1s, const float *__restrict__ sigma_p,
                           const uint8_t *__restrict__ bits, uint32_t *__restrict__ startidx,
                           uint32_t *__restrict__ counts, bool *__restrict__ fallback, float *__restrict__ xyzs,
                           float *__restrict__ dss, float *__restrict__ zs) {
    const int wave = (int)threadIdx.x / kWave;
    const uint32_t r = blockIdx.x * (kBlock / kWave) + wave;
    if (r >= n_rays) return;
    place_samples(sa, r, sigma_p[(size_t)r * kBins + lane_id()], t0s[r], t1s[r], rays_o, rays_d, bits, startidx,
                  counts, fallback, xyzs, dss, zs);
}

// The proposal's density at ray r's bin lane: the 5-level encoding of the bin
// centre and the MLP, in registers (weights w and level tables lt staged in LDS).
template <typename TT>
__device__ __forceinline__ float bin_sigma(const HashGridDescriptor &g, const LevelTables &lt, const float *w,
                                           const TT *__restrict__ table, const float *__restrict__ rays_o,
                                           const float *__restrict__ rays_d, uint32_t r, float t0, float t1) {
    const int lane = lane_id();
    const float D = (t1 - t0) / (float)kBins;
    const float tc = t0 + ((float)lane + 0.5f) * D;
    const Divisor two_b = make_divisor(2.0f * g.bound);
    const float px = divide(two_b, rays_o[3 * (size_t)r + 0] + tc * rays_d[3 * (size_t)r + 0] + g.bound);
    const float py = divide(two_b, rays_o[3 * (size_t)r + 1] + tc * rays_d[3 * (size_t)r + 1] + g.bound);
    const float pz = divide(two_b, rays_o[3 * (size_t)r + 2] + tc * rays_d[3 * (size_t)r + 2] + g.bound);
    float x[kIn];
    const bool pow2 = (g.T & (g.T - 1u)) == 0;
#pragma unroll
    for (int l = 0; l < kIn / 2; ++l) {
        uint32_t idx[8];
        float wt[8];
        level_vertices(px, py, pz, lt.scale[l], lt.res[l], (uint32_t)l < g.first_hash_level, g.T, pow2, lt.offset[l],
                       g.n_rows - 1u, idx, wt);
        float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const TT *row = table + (size_t)idx[c] * 2;
            a0 = a0 + to_f32<TT>(row[0]) * wt[c];
            a1 = a1 + to_f32<TT>(row[1]) * wt[c];
        }
        x[2 * l] = a0;
        x[2 * l + 1] = a1;
    }
    float h[kHid];
    return t1 > t0 ? det_expf(hidden_and_logit(w, x, h)) : 0.0f;
}

template <typename PT>
__device__ __forceinline__ void stage_proposal(const HashGridDescriptor &g, float *w, LevelTables &lt,
                                               const PT *__restrict__ params) {
    if (threadIdx.x < kIn / 2) {
        lt.scale[threadIdx.x] = g.scales[threadIdx.x];
        lt.res[threadIdx.x] = g.resolutions[threadIdx.x];
        lt.offset[threadIdx.x] = g.offsets[threadIdx.x];
    }
    stage_params(w, params);  // ends with __syncthreads
}

// Fused: lane b evaluates the proposal (encoder + MLP) at its bin centre, then the
// ray's samples are placed as in proposal_sample.  No bin points, encodings or
// densities go through memory.
template <typename TT, typename PT>
__global__ void __launch_bounds__(kBlock)
    proposal_render_kernel(uint32_t n_rays, HashGridDescriptor g, SampleArgs sa, const float *__restrict__ rays_o,
                           const float *__restrict__ rays_d, const float *__restrict__ t0s,
                           const float *__restrict__ t1s, const TT *__restrict__ table, const PT *__restrict__ params,
                           const uint8_t *__restrict__ bits, uint32_t *__restrict__ startidx,
                           uint32_t *__restrict__ counts, bool *__restrict__ fallback, float *__restrict__ xyzs,
                           float *__restrict__ dss, float *__restrict__ zs) {
    __shared__ float w[kParams];
    __shared__ LevelTables lt;
    stage_proposal(g, w, lt, params);
    const int wave = (int)threadIdx.x / kWave;
    const uint32_t r = blockIdx.x * (kBlock / kWave) + wave;
    if (r >= n_rays) return;
    const float t0 = t0s[r], t1 = t1s[r];
    const float sigma_b = bin_sigma(g, lt, w, table, rays_o, rays_d, r, t0, t1);
    place_samples(sa, r, sigma_b, t0, t1, rays_o, rays_d, bits, startidx, counts, fallback, xyzs, dss, zs);
}

// ---------------------------------------------------------------------------
// Training samples placed by the proposal, in march_rays' output layout (rays'
// samples contiguous in ray order, invalid rays and the unused tail empty), so the
// rest of the training step runs unchanged on them.  Stratified: sample j of ray r
// sits at the (j + noise_r) / n_per_ray quantile.  Passes:
//   1. count: the kept samples per ray -> rays_n_samples, and each kept sample's t,
//      ds and position staged at [r * 64 + rank] of the caller's workspace;
//   2. scan (one workgroup): start = exclusive prefix in ray order; a ray is valid
//      iff t_end > t_start and its samples fit (start + n <= total_samples), as
//      march_rays decides; next_sample_write_location = every ray's count,
//      number_of_exceeded_samples = the counts of the rays that did not fit;
//   3. write: each valid ray's staged samples are copied to its start;
//   4. the samples past the valid ones are zeroed.
// ---------------------------------------------------------------------------

// Workspace (stage): five f32 planes of n_rays * 64, the t, ds, x, y and z of ray r's
// kept sample k at [r * 64 + k] of each.
template <typename TT, typename PT>
__global__ void __launch_bounds__(kBlock)
    proposal_train_kernel(uint32_t n_rays, HashGridDescriptor g, SampleArgs sa, const float *__restrict__ rays_o,
                          const float *__restrict__ rays_d, const float *__restrict__ t0s,
                          const float *__restrict__ t1s, const float *__restrict__ noises,
                          const TT *__restrict__ table, const PT *__restrict__ params, const uint8_t *__restrict__ bits,
                          uint32_t *__restrict__ rays_n_samples, float *__restrict__ stage) {
    __shared__ float w[kParams];
    __shared__ LevelTables lt;
    stage_proposal(g, w, lt, params);
    const int wave = (int)threadIdx.x / kWave, lane = lane_id();
    const uint32_t r = blockIdx.x * (kBlock / kWave) + wave;
    if (r >= n_rays) return;
    const float t0 = t0s[r], t1 = t1s[r];
    const float sigma_b = bin_sigma(g, lt, w, table, rays_o, rays_d, r, t0, t1);
    const SamplePlan pl = plan_samples(sa, r, sigma_b, t0, t1, noises[r], rays_o, rays_d, bits);
    const uint64_t kept = __ballot(pl.keep);
    if (pl.keep) {
        const size_t plane = (size_t)n_rays * kBins;
        const size_t at = (size_t)r * kBins + (uint32_t)__popcll(kept & ((1ull << lane) - 1ull));
        stage[at] = pl.t;
        stage[plane + at] = pl.ds;
#pragma unroll
        for (int c = 0; c < 3; ++c) stage[(2 + c) * plane + at] = pl.p[c];
    }
    if (lane == 0) rays_n_samples[r] = (uint32_t)__popcll(kept);
}

__global__ void __launch_bounds__(kBlock)
    proposal_train_write_kernel(uint32_t n_rays, const float *__restrict__ rays_d, const float *__restrict__ stage,
                                const uint32_t *__restrict__ rays_n_samples, const uint32_t *__restrict__ startidx,
                                const bool *__restrict__ valid, uint32_t *__restrict__ idcs, float *__restrict__ xyzs,
                                float *__restrict__ dirs, float *__restrict__ dss, float *__restrict__ zs) {
    const int wave = (int)threadIdx.x / kWave, lane = lane_id();
    const uint32_t r = blockIdx.x * (kBlock / kWave) + wave;
    if (r >= n_rays || !valid[r]) return;
    if ((uint32_t)lane >= rays_n_samples[r]) return;  // <= 64: a count of one ballot
    const size_t plane = (size_t)n_rays * kBins, at = (size_t)r * kBins + lane;
    const size_t j = (size_t)startidx[r] + lane;
    idcs[j] = r;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        xyzs[3 * j + c] = stage[(2 + c) * plane + at];
        dirs[3 * j + c] = rays_d[3 * (size_t)r + c];
    }
    dss[j] = stage[plane + at];
    zs[j] = stage[at];
}

// Rays in chunks of 8 x 1024: every load of a chunk is issued up front (coalesced: ray
// c + 1024 q + thread), then eight block scans in ray order place them.
constexpr int kTrainScanThreads = 1024, kTrainScanPer = 8;
__global__ void __launch_bounds__(kTrainScanThreads)
    proposal_train_scan_kernel(uint32_t n_rays, uint32_t total_samples, const float *__restrict__ t0s,
                               const float *__restrict__ t1s, uint32_t *__restrict__ next_loc,
                               uint32_t *__restrict__ exceeded, bool *__restrict__ valid,
                               uint32_t *__restrict__ rays_n_samples, uint32_t *__restrict__ startidx) {
    uint64_t carry = 0, over = 0;
    for (uint64_t c = 0; c < n_rays; c += (uint64_t)kTrainScanThreads * kTrainScanPer) {
        uint32_t n[kTrainScanPer];
        bool seg[kTrainScanPer];
#pragma unroll
        for (int q = 0; q < kTrainScanPer; ++q) {
            const uint64_t r = c + (uint64_t)q * kTrainScanThreads + threadIdx.x;
            n[q] = r < n_rays ? rays_n_samples[r] : 0u;
            seg[q] = r < n_rays && t1s[r] > t0s[r];
        }
        uint32_t miss = 0;
#pragma unroll
        for (int q = 0; q < kTrainScanPer; ++q) {
            const uint64_t r = c + (uint64_t)q * kTrainScanThreads + threadIdx.x;
            uint32_t tot;
            const uint32_t ex = block_exclusive_scan<kTrainScanThreads>(n[q], &tot);
            const uint64_t start = carry + ex;
            const bool ok = seg[q] && start < total_samples && start + n[q] <= total_samples;
            if (r < n_rays) {
                valid[r] = ok;
                rays_n_samples[r] = ok ? n[q] : 0u;
                startidx[r] = ok && n[q] > 0 ? (uint32_t)start : 0u;
                miss += ok ? 0u : n[q];
            }
            carry += tot;
        }
        uint32_t miss_tot;
        block_exclusive_scan<kTrainScanThreads>(miss, &miss_tot);
        over += miss_tot;
    }
    if (threadIdx.x == 0) {
        *next_loc = (uint32_t)min<uint64_t>(carry, 0xffffffffull);
        *exceeded = (uint32_t)min<uint64_t>(over, 0xffffffffull);
    }
}

// Zeroes samples [next_loc - exceeded, total_samples).
__global__ void __launch_bounds__(kBlock)
    proposal_train_tail_kernel(uint32_t total_samples, const uint32_t *__restrict__ next_loc,
                               const uint32_t *__restrict__ exceeded, uint32_t *__restrict__ idcs,
                               float *__restrict__ xyzs, float *__restrict__ dirs, float *__restrict__ dss,
                               float *__restrict__ zs) {
    const uint32_t measured = *next_loc - *exceeded;
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t j = (size_t)measured + blockIdx.x * kBlock + threadIdx.x; j < total_samples; j += stride) {
        idcs[j] = 0u;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xyzs[3 * j + c] = 0.0f;
            dirs[3 * j + c] = 0.0f;
        }
        dss[j] = 0.0f;
        zs[j] = 0.0f;
    }
}

__global__ void zero_f32_kernel(float *p) {
    if (threadIdx.x == 0) *p = 0.0f;
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void proposal_points(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "proposal_points";
    ProposalRaysDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.n_rays == 0) return;
    if (!check_buffers(buffers, 5, target)) return;
    const uint64_t n = (uint64_t)d.n_rays * kBins;
    hipLaunchKernelGGL(proposal_points_kernel, dim3((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, d.n_rays, (const float *)buffers[0], (const float *)buffers[1],
                       (const float *)buffers[2], (const float *)buffers[3], (float *)buffers[4]);
    check_launch(target);
}

static bool valid_mlp_desc(const ProposalMlpDescriptor &d, const char *target) {
    if (d.enc_layout > NGP_ENC_LEVEL_MAJOR || d.param_dtype > NGP_DTYPE_F16) {
        set_error(kBadDescriptor, "%s: enc_layout and param_dtype must be 0 or 1", target);
        return false;
    }
    return true;
}

extern "C" void proposal_mlp_forward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "proposal_mlp_forward";
    ProposalMlpDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!valid_mlp_desc(d, target)) return;
    if (d.n_points == 0) return;
    if (!check_buffers(buffers, 3, target)) return;
    const uint32_t grid = std::min<uint32_t>(div_up(d.n_points, kBlock), 2048);
    const bool lm = d.enc_layout == NGP_ENC_LEVEL_MAJOR;
    if (d.param_dtype == NGP_DTYPE_F16)
        hipLaunchKernelGGL(proposal_fwd_kernel<__half>, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, d.n_points,
                           lm, (const float *)buffers[0], (const __half *)buffers[1], (float *)buffers[2]);
    else
        hipLaunchKernelGGL(proposal_fwd_kernel<float>, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, d.n_points,
                           lm, (const float *)buffers[0], (const float *)buffers[1], (float *)buffers[2]);
    check_launch(target);
}

extern "C" void proposal_mlp_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "proposal_mlp_backward";
    ProposalMlpDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!valid_mlp_desc(d, target)) return;
    if (d.n_workspace_blocks == 0) {
        set_error(kBadDescriptor, "%s: n_workspace_blocks must be positive", target);
        return;
    }
    if (!check_buffer_list(buffers, {1, 4, 5}, target)) return;
    if (d.n_points > 0 && !check_buffers(buffers, 6, target)) return;
    hipStream_t s = (hipStream_t)stream;
    const bool lm = d.enc_layout == NGP_ENC_LEVEL_MAJOR;
    if (d.param_dtype == NGP_DTYPE_F16)
        hipLaunchKernelGGL(proposal_bwd_kernel<__half>, dim3(d.n_workspace_blocks), dim3(kBlock), 0, s, d.n_points, lm,
                           (const float *)buffers[0], (const __half *)buffers[1], (const float *)buffers[2],
                           (float *)buffers[3], (float *)buffers[5]);
    else
        hipLaunchKernelGGL(proposal_bwd_kernel<float>, dim3(d.n_workspace_blocks), dim3(kBlock), 0, s, d.n_points, lm,
                           (const float *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                           (float *)buffers[3], (float *)buffers[5]);
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(proposal_sum_rows_kernel, dim3(kParams), dim3(kBlock), 0, s,
                       d.n_workspace_blocks, (const float *)buffers[5], (float *)buffers[4]);
    check_launch(target);
}

extern "C" void proposal_loss(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "proposal_loss";
    ProposalLossDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!(d.eps > 0.0f)) {
        set_error(kBadDescriptor, "%s: eps must be positive", target);
        return;
    }
    if (!check_buffer_list(buffers, {11}, target)) return;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, s, (float *)buffers[11]);
    if (d.n_rays > 0) {
        if (!check_buffers(buffers, 12, target)) return;
        const uint32_t grid = div_up(d.n_rays, kBlock / kWave);
        hipLaunchKernelGGL(proposal_loss_kernel, dim3(grid), dim3(kBlock), 0, s, d.n_rays, d.eps,
                           (const uint32_t *)buffers[0], (const uint32_t *)buffers[1], (const float *)buffers[2],
                           (const float *)buffers[3], (const float *)buffers[4], (const float *)buffers[5],
                           (const float *)buffers[6], (const float *)buffers[7], (const bool *)buffers[8],
                           (const uint32_t *)buffers[9], (float *)buffers[10], (float *)buffers[11]);
    }
    check_launch(target);
}

extern "C" void proposal_sample(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "proposal_sample";
    ProposalSampleDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.n_per_ray == 0 || d.n_per_ray > kBins || !(d.mass_threshold >= 0.0f) ||
        d.mode > NGP_PROPOSAL_SUPPORT || !(d.eps >= 0.0f && d.eps < 0.5f) || !(d.empty_mass >= 0.0f)) {
        set_error(kBadDescriptor, "%s: n_per_ray must be in [1, %d], mass_threshold >= 0, mode 0 or 1, eps in "
                  "[0, 0.5), empty_mass >= 0", target, kBins);
        return;
    }
    const uint32_t G = d.G;
    const bool g_ok = G != 0 && G <= 1024 && (G & (G - 1)) == 0 && (uint64_t)d.K * G * G * G <= 0xffffffffull;
    if (d.diagonal_n_steps == 0 || d.K == 0 || !g_ok || !(d.bound > 0.0f) || !(d.stepsize_portion >= 0.0f)) {
        set_error(kBadDescriptor, "%s: invalid marching statics (need diagonal_n_steps, K > 0, G a power of two "
                  "<= 1024, bound > 0, portion >= 0)", target);
        return;
    }
    if (d.n_rays == 0) return;
    if (!check_buffers(buffers, 12, target)) return;
    SampleArgs sa;
    sa.n_per_ray = d.n_per_ray;
    sa.mode = d.mode;
    sa.mass_threshold = d.mass_threshold;
    sa.eps = d.eps;
    sa.empty_mass = d.empty_mass;
    sa.mp = make_march_params(d.diagonal_n_steps, d.K, d.G, d.bound, d.stepsize_portion);
    hipLaunchKernelGGL(proposal_sample_kernel, dim3(div_up(d.n_rays, kBlock / kWave)), dim3(kBlock), 0,
                       (hipStream_t)stream, d.n_rays, sa, (const float *)buffers[0], (const float *)buffers[1],
                       (const float *)buffers[2], (const float *)buffers[3], (const float *)buffers[4],
                       (const uint8_t *)buffers[5], (uint32_t *)buffers[6], (uint32_t *)buffers[7], (bool *)buffers[8],
                       (float *)buffers[9], (float *)buffers[10], (float *)buffers[11]);
    check_launch(target);
}

// Checks a ProposalRenderDescriptor (the proposal grid, placement and marching
// statics) and fills the sampler's arguments.
static bool check_render_descriptor(const ProposalRenderDescriptor &d, SampleArgs &sa, const char *target) {
    const HashGridDescriptor &g = d.grid;
    const ProposalSampleDescriptor &q = d.sample;
    if (g.L != (uint32_t)(kIn / 2) || g.F != 2 || g.T == 0 || g.n_rows == 0 || g.first_hash_level > g.L ||
        g.offsets[0] != 0 || g.offsets[g.L] != g.n_rows || !(g.bound > 0.0f) || g.table_dtype > NGP_DTYPE_F16 ||
        d.param_dtype > NGP_DTYPE_F16) {
        set_error(kBadDescriptor, "%s: needs the proposal grid (L = 5, F = 2) with a consistent level table", target);
        return false;
    }
    for (uint32_t l = 0; l < g.L; ++l) {
        if (g.offsets[l + 1] < g.offsets[l]) {
            set_error(kBadDescriptor, "%s: offsets must be non-decreasing", target);
            return false;
        }
    }
    const uint32_t G = q.G;
    const bool g_ok = G != 0 && G <= 1024 && (G & (G - 1)) == 0 && (uint64_t)q.K * G * G * G <= 0xffffffffull;
    if (q.n_per_ray == 0 || q.n_per_ray > kBins || !(q.mass_threshold >= 0.0f) || q.mode > NGP_PROPOSAL_SUPPORT ||
        !(q.eps >= 0.0f && q.eps < 0.5f) || q.diagonal_n_steps == 0 || q.K == 0 || !g_ok || !(q.bound > 0.0f) ||
        !(q.stepsize_portion >= 0.0f) || !(q.empty_mass >= 0.0f)) {
        set_error(kBadDescriptor, "%s: invalid sample placement or marching statics", target);
        return false;
    }
    sa.n_per_ray = q.n_per_ray;
    sa.mode = q.mode;
    sa.mass_threshold = q.mass_threshold;
    sa.eps = q.eps;
    sa.empty_mass = q.empty_mass;
    sa.mp = make_march_params(q.diagonal_n_steps, q.K, q.G, q.bound, q.stepsize_portion);
    return true;
}

extern "C" void proposal_render_samples(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "proposal_render_samples";
    ProposalRenderDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    const HashGridDescriptor &g = d.grid;
    const ProposalSampleDescriptor &q = d.sample;
    SampleArgs sa;
    if (!check_render_descriptor(d, sa, target)) return;
    if (q.n_rays == 0) return;
    if (!check_buffers(buffers, 13, target)) return;
    const dim3 grid(div_up(q.n_rays, kBlock / kWave));
    hipStream_t s = (hipStream_t)stream;
    const float *ro = (const float *)buffers[0], *rd = (const float *)buffers[1];
    const float *t0 = (const float *)buffers[2], *t1 = (const float *)buffers[3];
    const uint8_t *bits = (const uint8_t *)buffers[6];
    uint32_t *start = (uint32_t *)buffers[7], *cnt = (uint32_t *)buffers[8];
    bool *fb = (bool *)buffers[9];
    float *xyzs = (float *)buffers[10], *dss = (float *)buffers[11], *zs = (float *)buffers[12];
    const bool t16 = g.table_dtype == NGP_DTYPE_F16, p16 = d.param_dtype == NGP_DTYPE_F16;
    if (t16 && p16)
        hipLaunchKernelGGL((proposal_render_kernel<__half, __half>), grid, dim3(kBlock), 0, s, q.n_rays, g, sa, ro, rd,
                           t0, t1, (const __half *)buffers[4], (const __half *)buffers[5], bits, start, cnt, fb, xyzs,
                           dss, zs);
    else if (t16)
        hipLaunchKernelGGL((proposal_render_kernel<__half, float>), grid, dim3(kBlock), 0, s, q.n_rays, g, sa, ro, rd,
                           t0, t1, (const __half *)buffers[4], (const float *)buffers[5], bits, start, cnt, fb, xyzs,
                           dss, zs);
    else if (p16)
        hipLaunchKernelGGL((proposal_render_kernel<float, __half>), grid, dim3(kBlock), 0, s, q.n_rays, g, sa, ro, rd,
                           t0, t1, (const float *)buffers[4], (const __half *)buffers[5], bits, start, cnt, fb, xyzs,
                           dss, zs);
    else
        hipLaunchKernelGGL((proposal_render_kernel<float, float>), grid, dim3(kBlock), 0, s, q.n_rays, g, sa, ro, rd,
                           t0, t1, (const float *)buffers[4], (const float *)buffers[5], bits, start, cnt, fb, xyzs,
                           dss, zs);
    check_launch(target);
}

extern "C" void proposal_train_samples(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "proposal_train_samples";
    ProposalTrainDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    const HashGridDescriptor &g = d.render.grid;
    const uint32_t n = d.render.sample.n_rays;
    SampleArgs sa;
    if (!check_render_descriptor(d.render, sa, target)) return;
    if (d.render.sample.mode != NGP_PROPOSAL_QUANTILES || d.total_samples == 0 ||
        d.n_uniform > (uint32_t)kBins - d.render.sample.n_per_ray) {
        set_error(kBadDescriptor, "%s: needs quantile placement, total_samples > 0 and n_per_ray + n_uniform <= %d",
                  target, kBins);
        return;
    }
    sa.n_uniform = d.n_uniform;
    if (!check_buffers(buffers, n > 0 ? 19 : 18, target)) return;
    hipStream_t s = (hipStream_t)stream;
    const float *ro = (const float *)buffers[0], *rd = (const float *)buffers[1];
    const float *t0 = (const float *)buffers[2], *t1 = (const float *)buffers[3], *noise = (const float *)buffers[4];
    const uint8_t *bits = (const uint8_t *)buffers[7];
    uint32_t *next_loc = (uint32_t *)buffers[8], *exceeded = (uint32_t *)buffers[9];
    bool *valid = (bool *)buffers[10];
    uint32_t *ns = (uint32_t *)buffers[11], *start = (uint32_t *)buffers[12], *idcs = (uint32_t *)buffers[13];
    float *xyzs = (float *)buffers[14], *dirs = (float *)buffers[15], *dss = (float *)buffers[16];
    float *zs = (float *)buffers[17];
    const bool t16 = g.table_dtype == NGP_DTYPE_F16, p16 = d.render.param_dtype == NGP_DTYPE_F16;
    const dim3 grid(std::max<uint32_t>(1, div_up(n, kBlock / kWave)));
    if (n > 0) {
        float *stage = (float *)buffers[18];
#define NGP_TRAIN_LAUNCH(TT, PT)                                                                                   \
    hipLaunchKernelGGL((proposal_train_kernel<TT, PT>), grid, dim3(kBlock), 0, s, n, g, sa, ro, rd, t0, t1, noise, \
                       (const TT *)buffers[5], (const PT *)buffers[6], bits, ns, stage)
        if (t16 && p16) NGP_TRAIN_LAUNCH(__half, __half);
        else if (t16) NGP_TRAIN_LAUNCH(__half, float);
        else if (p16) NGP_TRAIN_LAUNCH(float, __half);
        else NGP_TRAIN_LAUNCH(float, float);
#undef NGP_TRAIN_LAUNCH
        if (!check_launch(target)) return;
        hipLaunchKernelGGL(proposal_train_scan_kernel, dim3(1), dim3(kTrainScanThreads), 0, s, n, d.total_samples, t0,
                           t1, next_loc, exceeded, valid, ns, start);
        if (!check_launch(target)) return;
        hipLaunchKernelGGL(proposal_train_write_kernel, grid, dim3(kBlock), 0, s, n, rd, stage, ns, start, valid, idcs,
                           xyzs, dirs, dss, zs);
        if (!check_launch(target)) return;
    } else {
        hipLaunchKernelGGL(proposal_train_scan_kernel, dim3(1), dim3(kTrainScanThreads), 0, s, n, d.total_samples, t0,
                           t1, next_loc, exceeded, valid, ns, start);
        if (!check_launch(target)) return;
    }
    const uint32_t tail_grid = std::max<uint32_t>(1, std::min<uint32_t>(div_up(d.total_samples, kBlock), 1024));
    hipLaunchKernelGGL(proposal_train_tail_kernel, dim3(tail_grid), dim3(kBlock), 0, s, d.total_samples, next_loc,
                       exceeded, idcs, xyzs, dirs, dss, zs);
    check_launch(target);
}
