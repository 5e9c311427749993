// march.hip — occupancy-grid ray marching for training and inference.
//
// march_rays:            deps/volume-rendering-jax/lib/impl/marching.cu:101-268 (kernel), 435-522 (launcher)
// march_rays_inference:  marching.cu:271-397 (kernel), 524-604 (launcher)
//
// Training marcher.  The reference reserves each ray's output range with an
// atomicAdd in scheduling order and marches every ray twice.  Here the ranges are
// an exclusive prefix sum of the per-ray counts in ray order, which makes the
// output deterministic and equal to processing rays one after another:
//   1. march_count:  one wavefront per ray counts its samples; the block total is
//                    parked in rays_sample_startidx[first ray of the block] (slot
//                    owned by the block, so no scratch memory).  Rays are counted
//                    in ray-ordered launches (2048 rays, up to 16384, up to
//                    131072, the rest); a later launch's workgroups first add the
//                    parked totals of the earlier launches, and a launch that finds
//                    the sample buffer already full skips marching, which changes
//                    no output (its rays are invalid regardless of their counts).
//                    The first launch's rays also record their samples' t in the
//                    sample outputs, which are not written yet (a "t record" of up
//                    to `slot` floats per ray);
//   2. march_scan:   one workgroup scans the block totals, finds the ray where the
//                    sample buffer fills up and writes the two helper counters;
//   3. march_emit:   one wavefront per valid ray writes its samples' t to z_vals at
//                    its offset: copied from its t record (the first launch's rays,
//                    in training the whole batch), else marched again;
//   4. march_fill:   one wavefront per valid ray writes the outputs that follow from
//                    t (index, position, direction, step size);
//   5. march_finish: validity, final counts and offsets, zeros in the unused tail.
// A ray is valid iff the samples of the rays before it leave room (prefix <
// total_samples), t_end > t_start, and its own samples fit; rays after the buffer
// fills are invalid, as the reference's early exit at marching.cu:135 makes them.
//
// Wavefront marching.  Every t the reference's loop visits (marching.cu:162-192)
// lies on the sequence c_0 = t_start, c_{k+1} = c_k + step_size(c_k): an occupied
// step moves to the next element, an empty-cell skip to the first later element
// at or past the cell's exit.  So a wavefront generates 64 consecutive elements
// (one per lane, a serial chain of adds), looks up all 64 cells and exit distances
// at once, and walks the visited elements with ballots: runs of occupied lanes
// are emitted together, a skip is a search for the first lane past the exit.
// Samples go to their slots by lane prefix counts.  The walk visits exactly the
// reference's t values in its order; where the chain stops growing (t + ds == t)
// the ray finishes on one lane with the serial loop, which handles that case.
#include "ngp_common.h"
#include "ngp_device.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;
constexpr int kRaysPerBlock = kBlock / kWave;  // one wavefront per ray
constexpr int kScanThreads = 1024;
// parked total of a block whose launch found the sample buffer already full
constexpr uint32_t kPastFill = 0xffffffffu;

struct Ray {
    float ox, oy, oz, dx, dy, dz, idx, idy, idz, t1;
};

__device__ __forceinline__ Ray load_ray(const float *__restrict__ rays_o, const float *__restrict__ rays_d, uint32_t i,
                                        float t1) {
    Ray r;
    r.ox = rays_o[3 * (size_t)i];
    r.oy = rays_o[3 * (size_t)i + 1];
    r.oz = rays_o[3 * (size_t)i + 2];
    r.dx = rays_d[3 * (size_t)i];
    r.dy = rays_d[3 * (size_t)i + 1];
    r.dz = rays_d[3 * (size_t)i + 2];
    r.idx = 1.0f / r.dx;
    r.idy = 1.0f / r.dy;
    r.idz = 1.0f / r.dz;
    r.t1 = t1;
    return r;
}

// Where a marching wavefront records its samples: the t of sample k goes to ts[k]
// for k < cap (the other outputs follow from t, see march_fill_kernel).
struct RaySamples {
    float *ts;
    uint32_t cap;
};

__device__ __forceinline__ void put_t(const RaySamples &o, uint32_t k, float t) {
    if (k < o.cap) o.ts[k] = t;
}

// The t records of the first count launch's rays, kept in the sample outputs before
// they are written (xyzs, dirs, dss, then idcs, 8 floats per sample slot in all):
// ray i of that launch owns `slot` floats of the region q with
// first[q] <= i < first[q + 1].  A ray with more than `slot` samples, or past
// first[4], is marched again by the emit pass.
struct TStore {
    float *region[4];
    uint32_t first[5];
    uint32_t slot;
};

__device__ __forceinline__ float *t_record(const TStore &st, uint32_t i) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (i >= st.first[q] && i < st.first[q + 1]) return st.region[q] + (size_t)(i - st.first[q]) * st.slot;
    return nullptr;
}

// The reference's loop (marching.cu:162-192) from position t with n samples taken,
// on one lane.
template <bool kWrite>
__device__ uint32_t march_serial(const MarchParams &p, const uint8_t *__restrict__ bits, BitCache &cache, const Ray &r,
                                 float t, uint32_t n, uint32_t cap, const RaySamples &o) {
    while (n < cap && t < r.t1) {
        const float px = r.ox + t * r.dx, py = r.oy + t * r.dy, pz = r.oz + t * r.dz;
        const float ds = step_size(p, t);
        float gx, gy, gz, mb;
        if (lookup_cell(p, bits, cache, px, py, pz, ds, gx, gy, gz, mb)) {
            if (kWrite) put_t(o, n, t);
            ++n;
            t = t + ds;
        } else {
            const float next_t = cell_exit_t(p, t, px, py, pz, gx, gy, gz, mb, r.dx, r.dy, r.dz, r.idx, r.idy, r.idz);
            t = skip_to(p, t, ds, next_t);
        }
    }
    return n;
}

// Lane j's element of the chain c_0 = base, c_{j+1} = c_j + k (k > 0 constant).
// While the chain stays in base's binade, where every float is a multiple of the
// ulp u, each step adds the same m = k rounded to a multiple of u, so
// c_j = base + (j*m)*u with exact arithmetic.  Every lane then checks the
// recurrence c_{j+1} == c_j + k against its neighbour; the values are used only
// if all checks pass (by induction they are the chain), which fails where the
// chain leaves the binade, at rounding ties and for non-positive or non-normal
// base.  On success also returns c_64 in `next`; a chain that stops growing
// (next == c_63) is left to the serial generation, which detects it.
__device__ __forceinline__ bool chain_in_binade(float base, float k, int lane, float &w, float &next) {
    const uint32_t e8 = __float_as_uint(base) >> 23;  // sign bit included: negative base fails
    if (e8 < 24u || e8 >= 255u) return false;
    const float u = __uint_as_float((e8 - 23u) << 23);
    const float m = rintf(k / u);  // division by a power of two: exact
    if (!(m >= 1.0f && m <= 65536.0f)) return false;
    const float x = base + ((float)lane * m) * u;
    const float nx = x + k;
    const float up = __shfl_down(x, 1, kWave);
    const bool ok = lane == kWave - 1 || (up == nx && up != x);
    if (__ballot(!ok) != 0ull) return false;
    const float n63 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nx), kWave - 1));
    const float x63 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), kWave - 1));
    if (n63 == x63) return false;
    w = x;
    next = n63;
    return true;
}

// Lane j's element of the window starting at `base` (c_j of the chain from base) and
// the element after the window (c_64); `stall` where the chain stops growing.
__device__ __forceinline__ void chain_window(const MarchParams &p, bool const_step, float k_step, float base, int lane,
                                             float &w, float &next, bool &stall) {
    w = base;
    next = base;
    stall = false;
    if (const_step && chain_in_binade(base, k_step, lane, w, next)) return;  // 64 elements at once
    float c = base;
    if (const_step) {
#pragma unroll
        for (int q = 0; q < kWave; ++q) {
            if (lane == q) w = c;
            const float nx = c + k_step;
            stall |= nx == c;
            c = nx;
        }
    } else {
#pragma unroll 8
        for (int q = 0; q < kWave; ++q) {
            if (lane == q) w = c;
            const float nx = c + step_size(p, c);
            stall |= nx == c;
            c = nx;
        }
    }
    next = c;
}

// Windows of 64 chain elements a marching wavefront looks up together: their
// bitfield loads are all in flight before the first window is walked.
constexpr int kMarchWindows = 4;

// Counts the samples of one ray with the whole wavefront, starting at t_start,
// taking at most cap samples (kWrite: also records their t).  Call with all 64
// lanes active.
template <bool kWrite>
__device__ uint32_t march_ray_wave(const MarchParams &p, const uint8_t *__restrict__ bits, bool word_ok, const Ray &r,
                                   float t_start, uint32_t cap, const RaySamples &o) {
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    // with stepsize_portion 0 every step is the same constant, for every t
    const bool const_step = p.portion == 0.0f;
    const float k_step = step_size(p, 0.0f);
    const uint64_t *__restrict__ words = reinterpret_cast<const uint64_t *>(bits);
    uint32_t n = 0;
    float base = t_start;  // sequence element of lane 0 of the next window
    bool pending = false;  // a skip is searching for the first element >= thr
    float thr = 0.0f;
    while (n < cap) {
        // up to kMarchWindows windows: stop after a window that reaches t_end, and
        // before one whose chain stalls (finished on one lane below)
        float w[kMarchWindows];
        int nw = 0;
        bool stall = false;
#pragma unroll
        for (int q = 0; q < kMarchWindows; ++q) {
            if (nw == q && !stall) {
                float wq, next;
                bool st;
                chain_window(p, const_step, k_step, base, lane, wq, next, st);
                if (__ballot(st) != 0ull) {
                    stall = true;
                } else {
                    w[q] = wq;
                    nw = q + 1;
                    base = next;
                    // the window's last element at or past t_end: later windows are dead
                    if (!(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(wq), kWave - 1)) < r.t1)) break;
                }
            }
        }
        // cells of every window, the bitfield loads issued together
        uint32_t idx[kMarchWindows];
        float gx[kMarchWindows], gy[kMarchWindows], gz[kMarchWindows], mb[kMarchWindows];
        bool live[kMarchWindows], occ[kMarchWindows];
#pragma unroll
        for (int q = 0; q < kMarchWindows; ++q) {
            live[q] = q < nw && w[q] < r.t1;
            idx[q] = 0u;
            if (live[q]) {
                const float px = r.ox + w[q] * r.dx, py = r.oy + w[q] * r.dy, pz = r.oz + w[q] * r.dz;
                idx[q] = cell_index(p, px, py, pz, step_size(p, w[q]), gx[q], gy[q], gz[q], mb[q]);
            }
        }
        if (word_ok) {
            uint64_t wd[kMarchWindows];
#pragma unroll
            for (int q = 0; q < kMarchWindows; ++q) wd[q] = live[q] ? words[idx[q] >> 6] : 0ull;
#pragma unroll
            for (int q = 0; q < kMarchWindows; ++q) occ[q] = live[q] && ((wd[q] >> (idx[q] & 63u)) & 1ull);
        } else {
            uint32_t by[kMarchWindows];
#pragma unroll
            for (int q = 0; q < kMarchWindows; ++q) by[q] = live[q] ? bits[idx[q] >> 3] : 0u;
#pragma unroll
            for (int q = 0; q < kMarchWindows; ++q) occ[q] = live[q] && ((by[q] >> (idx[q] & 7u)) & 1u);
        }
        bool done = false;
#pragma unroll
        for (int q = 0; q < kMarchWindows; ++q) {
            if (q < nw && !done) {
                float ex = 0.0f;
                if (live[q] && !occ[q]) {
                    const float px = r.ox + w[q] * r.dx, py = r.oy + w[q] * r.dy, pz = r.oz + w[q] * r.dz;
                    ex = cell_exit_t(p, w[q], px, py, pz, gx[q], gy[q], gz[q], mb[q], r.dx, r.dy, r.dz, r.idx, r.idy,
                                     r.idz);
                }
                const uint64_t live_m = __ballot(live[q]);
                const uint64_t occ_m = __ballot(occ[q]);
                const uint32_t n0 = n;
                uint64_t emit = 0ull;
                int cur = 0;
                while (cur < kWave) {
                    if (pending) {
                        // skip_to stops at the first element with !(t < thr)
                        const uint64_t m = __ballot(!(w[q] < thr)) & (~0ull << cur);
                        if (m == 0ull) {
                            // the target lies past this window; if the window already
                            // reached t_end, so does the target and the ray is done
                            done = ((live_m >> (kWave - 1)) & 1ull) == 0ull;
                            break;
                        }
                        cur = __builtin_ctzll(m);
                        pending = false;
                    }
                    if (((live_m >> cur) & 1ull) == 0ull) {
                        done = true;
                        break;
                    }
                    if ((occ_m >> cur) & 1ull) {
                        const uint64_t from = occ_m >> cur;
                        const int run = ~from == 0ull ? kWave - cur : __builtin_ctzll(~from);
                        const uint32_t take = min((uint32_t)run, cap - n);
                        emit |= (take >= 64u ? ~0ull : ((1ull << take) - 1ull)) << cur;
                        n += take;
                        if (n >= cap) {
                            done = true;
                            break;
                        }
                        cur += run;
                    } else {
                        thr = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ex), cur));
                        pending = true;
                        cur += 1;
                    }
                }
                if (kWrite && ((emit >> lane) & 1ull)) put_t(o, n0 + (uint32_t)__popcll(emit & below), w[q]);
            }
        }
        if (done) break;
        if (stall) {
            // the chain stops growing at or after `base`: the serial loop on one lane
            uint32_t nn = n;
            if (lane == 0) {
                BitCache cache = make_bit_cache(bits, word_ok);
                float t = base;
                if (pending) {
                    // rest of skip_to: t is the first step past the skipped element
                    while (t < thr) {
                        const float s2 = t + step_size(p, t);
                        if (s2 == t) {
                            t = __builtin_huge_valf();
                            break;
                        }
                        t = s2;
                    }
                }
                nn = march_serial<kWrite>(p, bits, cache, r, t, n, cap, o);
            }
            return (uint32_t)__builtin_amdgcn_readlane((int)nn, 0);
        }
    }
    return n;
}

__device__ __forceinline__ float start_t(const MarchParams &p, float t0, float noise) {
    return t0 + step_size(p, t0) * noise;  // marching.cu:159-160
}

// Whether the blocks before `block_begin` (earlier launches, finished) already took
// total_samples samples: the workgroup adds their parked totals 1024 at a time and
// stops as soon as the sum reaches total_samples or meets a block an earlier
// launch marked kPastFill (that launch saw the buffer full).  Once the batch fills
// in the first launch, as in training, that is one chunk per workgroup.
__device__ bool filled_before(uint32_t block_begin, uint32_t total_samples,
                              const uint32_t *__restrict__ block_totals) {
    constexpr int kPer = 4;
    __shared__ uint64_t part[kBlock / kWave];
    __shared__ int verdict;
    const int wave = (int)threadIdx.x / kWave;
    uint64_t run = 0;
    for (uint32_t c0 = 0; c0 < block_begin; c0 += kBlock * kPer) {
        uint64_t sum = 0;
        bool past = false;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const uint32_t b = c0 + k * kBlock + threadIdx.x;
            if (b < block_begin) {
                const uint32_t v = block_totals[(size_t)b * kRaysPerBlock];
                if (v == kPastFill) past = true;
                else sum += v;
            }
        }
        sum = wave_sum64(sum);
        const bool wave_past = __ballot(past) != 0ull;
        if (lane_id() == 0) part[wave] = wave_past ? ~0ull : sum;
        __syncthreads();
        if (threadIdx.x == 0) {
            bool full = false;
            for (int w = 0; w < kBlock / kWave; ++w) {
                if (part[w] == ~0ull) full = true;
                else run += part[w];
            }
            verdict = full || run >= total_samples;
        }
        __syncthreads();
        if (verdict) return true;
    }
    return false;
}

// Counts the rays of blocks [block_begin, block_end) (grid-stride).  A launch whose
// predecessors already took total_samples samples marks its blocks kPastFill
// without marching.  kRecord (the first launch): rays with a t record also record
// their samples' t there.
template <bool kRecord>
__global__ void __launch_bounds__(kBlock)
    march_count_kernel(MarchParams p, bool word_ok, uint32_t n_rays, uint32_t total_samples, uint32_t block_begin,
                       uint32_t block_end, const float *__restrict__ rays_o, const float *__restrict__ rays_d,
                       const float *__restrict__ t_starts, const float *__restrict__ t_ends,
                       const float *__restrict__ noises, const uint8_t *__restrict__ bits, TStore store,
                       uint32_t *__restrict__ rays_n_samples,
                       uint32_t *__restrict__ block_totals /* = rays_sample_startidx */) {
    if (block_begin > 0 && filled_before(block_begin, total_samples, block_totals)) {
        for (uint32_t b = block_begin + blockIdx.x * kBlock + threadIdx.x; b < block_end; b += gridDim.x * kBlock)
            block_totals[(size_t)b * kRaysPerBlock] = kPastFill;
        return;
    }
    __shared__ uint32_t wave_n[kRaysPerBlock];
    const int wave = (int)threadIdx.x / kWave;
    for (uint32_t blk = block_begin + blockIdx.x; blk < block_end; blk += gridDim.x) {
        const uint32_t i = blk * kRaysPerBlock + wave;
        uint32_t n = 0;
        if (i < n_rays) {
            const float t0 = t_starts[i], t1 = t_ends[i];
            if (t1 > t0) {
                const Ray r = load_ray(rays_o, rays_d, i, t1);
                const float ts = start_t(p, t0, noises[i]);
                float *rec = kRecord ? t_record(store, i) : nullptr;
                if (rec != nullptr) n = march_ray_wave<true>(p, bits, word_ok, r, ts, p.n_cap, RaySamples{rec, store.slot});
                else n = march_ray_wave<false>(p, bits, word_ok, r, ts, p.n_cap, RaySamples{});
            }
            if (lane_id() == 0) rays_n_samples[i] = n;
        }
        if (lane_id() == 0) wave_n[wave] = n;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t s = 0;
#pragma unroll
            for (int w = 0; w < kRaysPerBlock; ++w) s += wave_n[w];
            block_totals[(size_t)blk * kRaysPerBlock] = (uint32_t)min(s, (uint64_t)kPastFill - 1);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const uint32_t c = a + b;
    return c < a ? 0xffffffffu : c;
}

// Single workgroup.  Replaces each parked block total with the block's exclusive
// prefix (saturated: only its comparison with total_samples matters) and writes
// next_sample_write_location / number_of_exceeded_samples.  Blocks are read in
// tiles of 4096 with all loads of a tile in flight; the scan stops at the first
// block of a skipped launch (every later block was skipped too, and their
// kPastFill slots already mark their rays as past the filled buffer).
constexpr int kScanPerThread = 4;
__global__ void __launch_bounds__(kScanThreads)
    march_scan_kernel(uint32_t n_rays, uint32_t n_blocks, uint32_t total_samples,
                      const uint32_t *__restrict__ rays_n_samples, uint32_t *__restrict__ block_slots,
                      uint32_t *__restrict__ next_loc, uint32_t *__restrict__ exceeded) {
    constexpr uint32_t kTile = kScanThreads * kScanPerThread;
    __shared__ uint32_t first_past;
    __shared__ uint32_t fill_block;
    __shared__ uint32_t wave_tot[kScanThreads / kWave];
    const uint32_t tid = threadIdx.x;
    const int lane = lane_id(), wave = (int)tid / kWave;
    if (tid == 0) fill_block = 0xffffffffu;
    uint32_t carry = 0;  // saturated prefix of the tiles before
    uint32_t scanned = 0;
    for (uint32_t t0 = 0; t0 < n_blocks; t0 += kTile) {
        if (tid == 0) first_past = 0xffffffffu;
        uint32_t v[kScanPerThread];
#pragma unroll
        for (int k = 0; k < kScanPerThread; ++k) {
            const uint32_t b = t0 + tid * kScanPerThread + k;
            v[k] = b < n_blocks ? block_slots[(size_t)b * kRaysPerBlock] : kPastFill;
        }
        __syncthreads();
        // first skipped block of the tile: block indices grow with the thread and slot,
        // so the lowest lane holding one has the wave's first; one LDS atomic per wave
        // (not one per skipped block, which serialised on the one address)
        uint32_t my_first = 0xffffffffu;
#pragma unroll
        for (int k = kScanPerThread - 1; k >= 0; --k)
            if (v[k] == kPastFill) my_first = t0 + tid * kScanPerThread + k;
        const uint64_t has = __ballot(my_first != 0xffffffffu);
        if (has != 0ull && lane == __ffsll((unsigned long long)has) - 1) atomicMin(&first_past, my_first);
        __syncthreads();
        const uint32_t stop = first_past;
        uint32_t mine = 0;
#pragma unroll
        for (int k = 0; k < kScanPerThread; ++k) {
            if (t0 + tid * kScanPerThread + k >= stop) v[k] = 0;
            mine = sat_add(mine, v[k]);
        }
        // inclusive saturated scan over the wave, then over the waves
        uint32_t incl = mine;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, kWave);
            if (lane >= d) incl = sat_add(incl, o);
        }
        if (lane == kWave - 1) wave_tot[wave] = incl;
        __syncthreads();
        uint32_t before = carry, tile_total = carry;
        for (int w = 0; w < kScanThreads / kWave; ++w) {
            if (w < wave) before = sat_add(before, wave_tot[w]);
            tile_total = sat_add(tile_total, wave_tot[w]);
        }
        uint32_t run = sat_add(before, incl - mine);  // exclusive prefix of this thread's first block
#pragma unroll
        for (int k = 0; k < kScanPerThread; ++k) {
            const uint32_t b = t0 + tid * kScanPerThread + k;
            if (b < stop && b < n_blocks) {
                // the first block whose inclusive prefix passes total_samples holds the fill point
                const uint32_t incl_b = sat_add(run, v[k]);
                if (incl_b > total_samples && run <= total_samples) atomicMin(&fill_block, b);
                block_slots[(size_t)b * kRaysPerBlock] = run;
                run = incl_b;
            }
        }
        carry = tile_total;
        __syncthreads();  // wave_tot and first_past are rewritten by the next tile
        if (stop != 0xffffffffu) {
            scanned = min(stop, n_blocks);
            break;
        }
        scanned = min(n_blocks, t0 + kTile);
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t grand = carry;
        if (grand <= total_samples) {
            *next_loc = grand;
            *exceeded = 0;
        } else {
            // walk the fill block's rays: first ray whose inclusive prefix passes total_samples
            const uint32_t b = fill_block < scanned ? fill_block : scanned - 1;  // set whenever grand > S
            uint64_t pre = block_slots[(size_t)b * kRaysPerBlock];
            const uint32_t first = b * kRaysPerBlock, last = min(n_rays, first + kRaysPerBlock);
            uint32_t measured = (uint32_t)pre, over = 0;
            for (uint32_t r = first; r < last; ++r) {
                const uint32_t n = rays_n_samples[r];
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

// Sample range of ray i from the scanned block prefixes (slot of i's block) and
// the counts of the rays before i in its block.
__device__ __forceinline__ uint64_t ray_start(const uint32_t *__restrict__ rays_n_samples,
                                              const uint32_t *__restrict__ block_slots, uint32_t i) {
    const uint32_t first = i / kRaysPerBlock * kRaysPerBlock;
    uint64_t start = block_slots[first];
    for (uint32_t j = first; j < i; ++j) start += rays_n_samples[j];
    return start;
}

// Wavefront per ray, grid-stride over rays: writes the t of every valid ray's
// samples to z_vals at the ray's range, copied from its t record, or marched again
// where it has none.  Stops at the first ray whose block prefix reaches
// total_samples (every later ray is past the filled buffer).  Runs before
// march_finish_kernel overwrites the block slots.
__global__ void __launch_bounds__(kBlock)
    march_emit_kernel(MarchParams p, bool word_ok, uint32_t n_rays, uint32_t total_samples,
                      const float *__restrict__ rays_o, const float *__restrict__ rays_d,
                      const float *__restrict__ t_starts, const float *__restrict__ t_ends,
                      const float *__restrict__ noises, const uint8_t *__restrict__ bits, TStore store,
                      uint32_t n_recorded, const uint32_t *__restrict__ rays_n_samples,
                      const uint32_t *__restrict__ block_slots, float *__restrict__ zs) {
    const uint32_t n_waves = gridDim.x * kRaysPerBlock;
    for (uint32_t i = blockIdx.x * kRaysPerBlock + threadIdx.x / kWave; i < n_rays; i += n_waves) {
        if (block_slots[i / kRaysPerBlock * kRaysPerBlock] >= total_samples) break;
        const uint64_t start = ray_start(rays_n_samples, block_slots, i);
        const uint32_t n = rays_n_samples[i];
        const float t0 = t_starts[i], t1 = t_ends[i];
        if (start < total_samples && t1 > t0 && start + n <= total_samples && n > 0) {
            const uint32_t s = (uint32_t)start;
            const float *rec = i < n_recorded && n <= store.slot ? t_record(store, i) : nullptr;
            if (rec != nullptr) {
                for (uint32_t k = lane_id(); k < n; k += kWave) zs[s + k] = rec[k];
            } else {
                const Ray r = load_ray(rays_o, rays_d, i, t1);
                march_ray_wave<true>(p, bits, word_ok, r, start_t(p, t0, noises[i]), min(p.n_cap, n),
                                     RaySamples{zs + s, 0xffffffffu});
            }
        }
    }
}

// Wavefront per valid ray, lanes over its samples: the outputs that follow from t
// (marching.cu:178-187): ray index, position o + t d, direction, step size.
__global__ void __launch_bounds__(kBlock)
    march_fill_kernel(MarchParams p, uint32_t n_rays, uint32_t total_samples, const float *__restrict__ rays_o,
                      const float *__restrict__ rays_d, const float *__restrict__ t_starts,
                      const float *__restrict__ t_ends, const uint32_t *__restrict__ rays_n_samples,
                      const uint32_t *__restrict__ block_slots, const float *__restrict__ zs,
                      uint32_t *__restrict__ idcs, float *__restrict__ xyzs, float *__restrict__ dirs,
                      float *__restrict__ dss) {
    const uint32_t n_waves = gridDim.x * kRaysPerBlock;
    for (uint32_t i = blockIdx.x * kRaysPerBlock + threadIdx.x / kWave; i < n_rays; i += n_waves) {
        if (block_slots[i / kRaysPerBlock * kRaysPerBlock] >= total_samples) break;
        const uint64_t start = ray_start(rays_n_samples, block_slots, i);
        const uint32_t n = rays_n_samples[i];
        const float t0 = t_starts[i], t1 = t_ends[i];
        if (start < total_samples && t1 > t0 && start + n <= total_samples && n > 0) {
            const Ray r = load_ray(rays_o, rays_d, i, t1);
            for (uint32_t k = lane_id(); k < n; k += kWave) {
                const size_t j = start + k;
                const float t = zs[j];
                idcs[j] = i;
                xyzs[3 * j + 0] = r.ox + t * r.dx;
                xyzs[3 * j + 1] = r.oy + t * r.dy;
                xyzs[3 * j + 2] = r.oz + t * r.dz;
                dirs[3 * j + 0] = r.dx;
                dirs[3 * j + 1] = r.dy;
                dirs[3 * j + 2] = r.dz;
                dss[j] = step_size(p, t);
            }
        }
    }
}

// Thread per ray: validity, final counts and start indices (replacing the block
// slots), and zeros in the sample slots no valid ray owns, [measured, total_samples).
__global__ void __launch_bounds__(kBlock)
    march_finish_kernel(uint32_t n_rays, uint32_t total_samples, const float *__restrict__ t_starts,
                        const float *__restrict__ t_ends, const uint32_t *__restrict__ next_loc,
                        const uint32_t *__restrict__ exceeded, bool *__restrict__ ray_is_valid,
                        uint32_t *__restrict__ rays_n_samples, uint32_t *__restrict__ rays_sample_startidx,
                        uint32_t *__restrict__ idcs, float *__restrict__ xyzs, float *__restrict__ dirs,
                        float *__restrict__ dss, float *__restrict__ zs) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const int lane = lane_id();
    // a block's rays are kRaysPerBlock consecutive lanes of one wavefront
    const uint32_t n = i < n_rays ? rays_n_samples[i] : 0u;
    const uint32_t slot = i < n_rays ? rays_sample_startidx[i / kRaysPerBlock * kRaysPerBlock] : 0u;
    const int k = lane % kRaysPerBlock;
    uint64_t start = slot;
#pragma unroll
    for (int d = 1; d < kRaysPerBlock; ++d) {
        const uint32_t o = __shfl_up(n, d, kWave);
        if (k >= d) start += o;
    }
    if (i < n_rays) {
        const float t0 = t_starts[i], t1 = t_ends[i];
        const bool valid = start < total_samples && t1 > t0 && start + n <= total_samples;
        ray_is_valid[i] = valid;
        rays_n_samples[i] = valid ? n : 0u;
        rays_sample_startidx[i] = (valid && n > 0) ? (uint32_t)start : 0u;
    }
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

// Terminated slots take new ray indices in slot order (the reference hands them
// out with atomicAdd, marching.cu:300), in three launches without scratch memory:
// per-block counts of terminated slots are parked in the first slot of each
// block's range of indices_out, one workgroup turns them into exclusive prefixes
// (and writes next_out), and every block then assigns its range.
constexpr int kAssignBlock = 1024;

__global__ void __launch_bounds__(kAssignBlock)
    assign_count_kernel(uint32_t n_rays, const bool *__restrict__ terminated, uint32_t *__restrict__ indices_out) {
    __shared__ uint32_t wave_cnt[kAssignBlock / kWave];
    const uint32_t j = blockIdx.x * kAssignBlock + threadIdx.x;
    const bool t = j < n_rays && terminated[j];
    const uint32_t c = (uint32_t)__popcll(__ballot(t));
    if (lane_id() == 0) wave_cnt[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < kAssignBlock / kWave; ++w) s += wave_cnt[w];
        indices_out[(size_t)blockIdx.x * kAssignBlock] = s;
    }
}

__global__ void __launch_bounds__(kScanThreads)
    assign_scan_kernel(uint32_t n_blocks, const uint32_t *__restrict__ next_in, uint32_t *__restrict__ next_out,
                       uint32_t *__restrict__ indices_out) {
    const uint32_t first = *next_in;  // read before anything is written (next_out may alias it)
    uint32_t carry = first;
    for (uint32_t c0 = 0; c0 < n_blocks; c0 += kScanThreads) {
        const uint32_t b = c0 + threadIdx.x;
        const uint32_t v = b < n_blocks ? indices_out[(size_t)b * kAssignBlock] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<kScanThreads>(v, &tot);
        if (b < n_blocks) indices_out[(size_t)b * kAssignBlock] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *next_out = carry;
}

__global__ void __launch_bounds__(kAssignBlock)
    assign_write_kernel(uint32_t n_rays, const bool *__restrict__ terminated, const uint32_t *__restrict__ indices_in,
                        uint32_t *__restrict__ indices_out) {
    __shared__ uint32_t base;
    __shared__ uint32_t wave_cnt[kAssignBlock / kWave];
    const uint32_t j = blockIdx.x * kAssignBlock + threadIdx.x;
    if (threadIdx.x == 0) base = indices_out[(size_t)blockIdx.x * kAssignBlock];
    const bool t = j < n_rays && terminated[j];
    const uint64_t m = __ballot(t);
    if (lane_id() == 0) wave_cnt[threadIdx.x / kWave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = base;
    const int wave = (int)threadIdx.x / kWave;
    for (int w = 0; w < wave; ++w) before += wave_cnt[w];
    const uint64_t below = (1ull << lane_id()) - 1ull;
    if (j < n_rays) indices_out[j] = t ? before + (uint32_t)__popcll(m & below) : indices_in[j];
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
    const uint32_t n_blocks = div_up(d.n_rays, kRaysPerBlock);
    // count in ray-ordered launches of 2048, 16384, 131072 and all rays; grids
    // of at most kMaxCountBlocks blocks (8192 resident wavefronts) stride over them
    constexpr uint32_t kMaxCountBlocks = 2048;
    const uint32_t bounds[5] = {0, std::min(n_blocks, 2048u / kRaysPerBlock),
                                std::min(n_blocks, 16384u / kRaysPerBlock), std::min(n_blocks, 131072u / kRaysPerBlock),
                                n_blocks};
    // t records of the first launch's rays in the sample outputs (8 floats per slot)
    const uint32_t n_first = std::min(d.n_rays, bounds[1] * kRaysPerBlock);
    TStore store{};
    store.slot = (uint32_t)std::min<uint64_t>(p.n_cap, 8ull * d.total_samples / n_first);
    {
        float *region[4] = {xyzs, dirs, dss, (float *)idcs};
        const uint64_t floats[4] = {3ull * d.total_samples, 3ull * d.total_samples, d.total_samples, d.total_samples};
        uint32_t first = 0;
        for (int q = 0; q < 4; ++q) {
            store.region[q] = region[q];
            store.first[q] = first;
            const uint64_t fit = store.slot ? floats[q] / store.slot : 0;
            first = (uint32_t)std::min<uint64_t>(n_first, first + fit);
        }
        store.first[4] = first;
    }
    for (int c = 0; c < 4; ++c) {
        if (bounds[c + 1] <= bounds[c]) continue;
        const dim3 grid(std::min(bounds[c + 1] - bounds[c], kMaxCountBlocks));
        if (c == 0)
            hipLaunchKernelGGL(march_count_kernel<true>, grid, dim3(kBlock), 0, stream, p, word_ok, d.n_rays,
                               d.total_samples, bounds[c], bounds[c + 1], rays_o, rays_d, t_starts, t_ends, noises,
                               bits, store, n_samples, startidx);
        else
            hipLaunchKernelGGL(march_count_kernel<false>, grid, dim3(kBlock), 0, stream, p, word_ok, d.n_rays,
                               d.total_samples, bounds[c], bounds[c + 1], rays_o, rays_d, t_starts, t_ends, noises,
                               bits, store, n_samples, startidx);
        if (!check_launch(target)) return;
    }
    hipLaunchKernelGGL(march_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, d.n_rays, n_blocks, d.total_samples,
                       n_samples, startidx, next_loc, exceeded);
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(march_emit_kernel, dim3(std::min(n_blocks, kMaxCountBlocks)), dim3(kBlock), 0, stream, p, word_ok,
                       d.n_rays, d.total_samples, rays_o, rays_d, t_starts, t_ends, noises, bits, store,
                       store.first[4], n_samples, startidx, zs);
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(march_fill_kernel, dim3(std::min(n_blocks, kMaxCountBlocks)), dim3(kBlock), 0, stream, p,
                       d.n_rays, d.total_samples, rays_o, rays_d, t_starts, t_ends, n_samples, startidx, zs, idcs, xyzs,
                       dirs, dss);
    if (!check_launch(target)) return;
    // a thread per ray, and enough workgroups for the tail's zeros when rays are few
    const uint32_t finish_grid = std::max(div_up(d.n_rays, kBlock), std::min(div_up(d.total_samples, kBlock), 2048u));
    hipLaunchKernelGGL(march_finish_kernel, dim3(finish_grid), dim3(kBlock), 0, stream, d.n_rays,
                       d.total_samples, t_starts, t_ends, next_loc, exceeded, valid, n_samples, startidx, idcs, xyzs,
                       dirs, dss, zs);
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
    const uint32_t n_assign = div_up(d.n_rays, kAssignBlock);
    hipLaunchKernelGGL(assign_count_kernel, dim3(n_assign), dim3(kAssignBlock), 0, stream, d.n_rays, terminated,
                       indices_out);
    hipLaunchKernelGGL(assign_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, n_assign, next_in, next_out,
                       indices_out);
    hipLaunchKernelGGL(assign_write_kernel, dim3(n_assign), dim3(kAssignBlock), 0, stream, d.n_rays, terminated,
                       indices_in, indices_out);
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(march_inference_kernel, dim3(div_up(d.n_rays, kBlock)), dim3(kBlock), 0, stream, p,
                       bitfield_words_ok(bits, d.K, d.G), d.n_total_rays, d.n_rays, d.march_steps_cap, rays_o, rays_d, t_starts, t_ends, bits,
                       indices_out, n_samples, t_out, xyzs, dss, zs);
    check_launch(target);
}
