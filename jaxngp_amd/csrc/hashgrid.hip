// hashgrid.hip — multiresolution hash encoding, forward and table gradient.
//
// Forward restates models/encoders.py:82-233 (HashGridEncoder.__call__): the
// reference evaluates it as XLA gather + lerp fusions and materialises [L,N,8]
// index and weight arrays; here each thread computes one (point, level) pair in
// registers: normalise, scale, floor, 8 vertex indices (dense x + y res + z res^2
// below first_hash_level, else x ^ y*2654435761 ^ z*805459861, both mod T, plus the
// level's row offset), trilinear weights, weighted sum of the gathered rows.
// Point-major output: threads are laid out level-fastest, so the L*F outputs of a
// point are written by L consecutive lanes as one contiguous run.  Level-major
// output: workgroups are level-affine per XCD (see level_block) so each level's
// gathers hit one L2.
//
// Backward is the VJP of the gather (encoders.py:226) and of the weighted sum
// (encoders.py:231): dtable[idx] += w * g, computed by a counting sort of the
// contributions into row slices that are summed in LDS (no global atomics).
//
// Arithmetic: f32, no contraction (matches the CPU oracle bit for bit on indices
// and weights).  Indices past the last row (only possible when the last level is
// dense) are clamped to the last row, as JAX clamps gather indices.
#include <hip/hip_fp16.h>

#include "ngp_common.h"
#include "hashgrid_device.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;

template <int F, typename TT>
__global__ void __launch_bounds__(kBlock)
    hashgrid_fwd_kernel(HashGridDescriptor d, const float *__restrict__ pos, const TT *__restrict__ table,
                        float *__restrict__ enc) {
    __shared__ LevelTables lt;
    if (threadIdx.x < d.L) {
        lt.scale[threadIdx.x] = d.scales[threadIdx.x];
        lt.res[threadIdx.x] = d.resolutions[threadIdx.x];
        lt.offset[threadIdx.x] = d.offsets[threadIdx.x];
    }
    __syncthreads();
    const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (gid >= (uint64_t)d.n_points * d.L) return;
    const uint32_t n = (uint32_t)(gid / d.L);
    const uint32_t l = (uint32_t)(gid - (uint64_t)n * d.L);
    const Divisor two_b = make_divisor(2.0f * d.bound);
    const float px = divide(two_b, pos[3 * (size_t)n + 0] + d.bound);
    const float py = divide(two_b, pos[3 * (size_t)n + 1] + d.bound);
    const float pz = divide(two_b, pos[3 * (size_t)n + 2] + d.bound);
    uint32_t idx[8];
    float w[8];
    level_vertices(px, py, pz, lt.scale[l], lt.res[l], l < d.first_hash_level, d.T, (d.T & (d.T - 1u)) == 0,
                   lt.offset[l], d.n_rows - 1u, idx, w);
    float acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const TT *row = table + (size_t)idx[c] * F;
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = acc[f] + to_f32<TT>(row[f]) * w[c];
    }
    float *out = enc + gid * F;
#pragma unroll
    for (int f = 0; f < F; ++f) out[f] = acc[f];
}

// Workgroups are dispatched round-robin over the 8 XCDs of the chip (block b runs
// on XCD b % 8), each with its own 4 MiB L2.  Level-affine kernels therefore give
// block b the level 8 * phase + b % 8: all blocks of one XCD work on the same level
// in a phase, so that level's rows (at most T * F * 4 B = 4 MiB for T = 2^19, F = 2)
// stay resident in that XCD's L2 for the gathers / atomics, instead of every L2
// streaming the whole 46.5 MiB table.
constexpr uint32_t kXcds = 8;
constexpr int kLmPointsPerThread = 2;
constexpr int kPointsPerBlock = kBlock * kLmPointsPerThread;

struct LevelBlock {
    uint32_t level, point0;
    bool live;
};

__device__ __forceinline__ LevelBlock level_block(uint32_t n_points, uint32_t L) {
    const uint32_t n_pg = (n_points + kPointsPerBlock - 1) / kPointsPerBlock;
    const uint32_t b = blockIdx.x;
    const uint32_t x = b % kXcds, rest = b / kXcds;
    const uint32_t pg = rest % n_pg, phase = rest / n_pg;
    LevelBlock lb;
    lb.level = phase * kXcds + x;
    lb.point0 = pg * kPointsPerBlock;
    lb.live = lb.level < L;
    return lb;
}

inline uint32_t level_grid(uint32_t n_points, uint32_t L) {
    const uint32_t n_pg = div_up(n_points, kPointsPerBlock);
    return kXcds * n_pg * div_up(L, kXcds);
}

// Level-major forward: each thread takes kLmPointsPerThread points of the block's
// level, kBlock apart, so the gathers of both are in flight together; output
// enc[(l * N + n) * F + f] written as one contiguous run per wavefront.  Points past
// the end load the last point and store nothing.
template <int F, typename TT>
__global__ void __launch_bounds__(kBlock)
    hashgrid_fwd_lm_kernel(HashGridDescriptor d, const float *__restrict__ pos, const TT *__restrict__ table,
                           float *__restrict__ enc) {
    constexpr int P = kLmPointsPerThread;
    const LevelBlock lb = level_block(d.n_points, d.L);
    if (!lb.live) return;
    const uint32_t l = lb.level;
    const Divisor two_b = make_divisor(2.0f * d.bound);
    uint32_t idx[P][8];
    float w[P][8];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const uint32_t n = lb.point0 + j * kBlock + threadIdx.x;
        const uint32_t np = n < d.n_points ? n : d.n_points - 1;
        const float px = divide(two_b, pos[3 * (size_t)np + 0] + d.bound);
        const float py = divide(two_b, pos[3 * (size_t)np + 1] + d.bound);
        const float pz = divide(two_b, pos[3 * (size_t)np + 2] + d.bound);
        level_vertices(px, py, pz, d.scales[l], d.resolutions[l], l < d.first_hash_level, d.T,
                       (d.T & (d.T - 1u)) == 0, d.offsets[l], d.n_rows - 1u, idx[j], w[j]);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        float acc[F];
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const TT *row = table + (size_t)idx[j][c] * F;
#pragma unroll
            for (int f = 0; f < F; ++f) acc[f] = acc[f] + to_f32<TT>(row[f]) * w[j][c];
        }
        const uint32_t n = lb.point0 + j * kBlock + threadIdx.x;
        if (n < d.n_points) {
            float *out = enc + ((size_t)l * d.n_points + n) * F;
#pragma unroll
            for (int f = 0; f < F; ++f) out[f] = acc[f];
        }
    }
}

// Backward without global atomics.  Device-scope atomics on this chip (float or
// integer, any scope) retire at ~21-27 G/s for the whole device on scattered
// addresses, two orders of magnitude below the gather rate, so contributions are
// routed to LDS by a deterministic counting sort over row slices of R rows
// (R rows of F floats fill 128 KiB of LDS):
//   1. count:   work item = (level, 2048-point chunk); LDS histogram of the item's
//               contributions per slice, stored slice-major;
//   2. scan:    exclusive scan of the counts -> each item's write offset per slice
//               (per-block sums, then each block adds the sums before its own);
//   3. scatter: the item recomputes its contributions (with dL_dencodings) and
//               writes (row within slice, values) entries at its offsets;
//   4. reduce:  workgroup s sums slice s's contiguous entries with LDS atomics and
//               stores its rows once (so no table zero-fill either).
// Within an item, lanes are 64 consecutive samples of one level; runs of equal rows
// (coarse levels: consecutive samples of a ray share cells) are summed with a
// segmented scan first, so one entry carries a run.
constexpr int kChunk = 2048;         // points per work item
constexpr int kItemThreads = 256;    // 8 points per thread per item
constexpr int kReduceThreads = 1024;
constexpr uint32_t kPart = 65536;    // entries one reduce workgroup sums
constexpr uint32_t kMaxSlices = 8192;
constexpr uint32_t kSliceBytes = 131072;
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;  // elements per thread in the scan kernels

// An entry is (row within slice, F values), padded to a power of two of words so
// that it is written and read with one or two vector accesses.
// Words of one sort entry: the row in its slice, then the F values, as f32 (F = 2:
// 12 bytes, one dwordx3 access) or, for f16 tables with F = 2 or 4, as f16 pairs
// (F = 2: 8 bytes) — the precision path accumulates its table gradient from
// f16-rounded contributions, as half-precision grids are trained in tiny-cuda-nn.
template <int F, bool kHalf>
constexpr int entry_words() {
    return kHalf && F == 2 ? 2 : (kHalf && F == 4 ? 3 : (F == 1 ? 2 : (F == 2 ? 3 : (F == 4 ? 8 : 16))));
}

inline int entry_words_host(uint32_t F, bool half) {
    return half && F == 2 ? 2 : (half && F == 4 ? 3 : (F == 1 ? 2 : (F == 2 ? 3 : (F == 4 ? 8 : 16))));
}

template <int F, bool kHalf>
__device__ __forceinline__ void pack_values(uint32_t *w, const float *v) {
    if constexpr (kHalf && (F == 2 || F == 4)) {
#pragma unroll
        for (int q = 0; q < F / 2; ++q) {
            const __half2 h = __floats2half2_rn(v[2 * q], v[2 * q + 1]);
            w[1 + q] = *reinterpret_cast<const uint32_t *>(&h);
        }
    } else {
#pragma unroll
        for (int f = 0; f < F; ++f) w[1 + f] = __float_as_uint(v[f]);
    }
}

template <int F, bool kHalf>
__device__ __forceinline__ float entry_value(const uint32_t *w, int f) {
    if constexpr (kHalf && (F == 2 || F == 4)) {
        const uint32_t u = w[1 + f / 2];
        const __half2 h = *reinterpret_cast<const __half2 *>(&u);
        return (f & 1) ? __high2float(h) : __low2float(h);
    } else {
        return __uint_as_float(w[1 + f]);
    }
}

// The last row level l's vertices can address.  A dense level's vertex
// coordinates reach res (a point on the upper face), so its rows run up to
// res + res^2 + res^3 past the level's offset, into the next level's rows, as the
// reference's indices do; hashed rows stay below offset + T.  Clamped to the table.
__host__ __device__ inline uint32_t level_last_row(const HashGridDescriptor &d, uint32_t l) {
    const uint64_t res = d.resolutions[l];
    uint64_t local = (uint64_t)d.T - 1u;
    if (l < d.first_hash_level) local = std::min<uint64_t>(local, res + res * res + res * res * res);
    return (uint32_t)std::min<uint64_t>((uint64_t)d.offsets[l] + local, (uint64_t)d.n_rows - 1u);
}

// The one-pass table gradient (F <= 2, every level's rows within kStageSlices
// slices): work item = (level, kBucketPoints consecutive points).
constexpr int kBucketThreads = 512;
constexpr int kBucketPhases = 2;
constexpr uint32_t kBucketPoints = kBucketThreads * kBucketPhases;
constexpr uint32_t kBucketEntries = kBucketPoints * 8;  // entries one item may write
constexpr uint32_t kStageSlices = 1024;                 // slices one level may span

struct BwdPlan {
    uint32_t R, n_slices, n_chunks, n_items;
    uint64_t M;             // n_slices * n_items counters
    uint64_t scan_blocks;   // blocks of the counter scan
    uint32_t max_parts;     // upper bound of the reduce grid
    uint64_t counts_off, sums_off, meta_off, parts_off, part_slice_off, partials_off, entries_off, bytes;
    // one-pass path: items of kBucketPoints points, each level's slices
    // [s_lo[l], s_hi[l]], per item and slice of its level the run's start and length
    bool bucket;
    uint32_t n_bchunks, n_bitems, span_stride;
    uint32_t s_lo[NGP_HASHGRID_MAX_LEVELS], s_hi[NGP_HASHGRID_MAX_LEVELS];
    uint64_t run_cnt_off, run_off_off, slice_np_off;
};

// Rows of one slice: R rows of F f64 accumulators fill kSliceBytes of LDS.  A power
// of two, so the kernels split a row into (slice, row in slice) with shifts.
template <int F>
constexpr uint32_t slice_rows() {
    static_assert((kSliceBytes / (8 * F) & (kSliceBytes / (8 * F) - 1)) == 0, "slice rows: a power of two");
    return kSliceBytes / (8 * F);
}

inline BwdPlan make_plan(const HashGridDescriptor &d) {
    auto align = [](uint64_t x) { return (x + 255) / 256 * 256; };
    BwdPlan p{};
    p.R = kSliceBytes / (8 * d.F);  // slice_rows<F>()
    p.n_slices = div_up(d.n_rows, p.R);
    const int ew = entry_words_host(d.F, d.table_dtype == NGP_DTYPE_F16);
    p.bucket = d.F <= 2;
    p.span_stride = 1;
    for (uint32_t l = 0; l < d.L; ++l) {
        p.s_lo[l] = d.offsets[l] / p.R;
        p.s_hi[l] = level_last_row(d, l) / p.R;
        const uint32_t span = p.s_hi[l] - p.s_lo[l] + 1u;
        if (span > kStageSlices) p.bucket = false;
        p.span_stride = std::max(p.span_stride, span);
    }
    if (p.bucket) {
        p.n_bchunks = std::max<uint32_t>(1, div_up(d.n_points, kBucketPoints));
        p.n_bitems = d.L * p.n_bchunks;
        const uint64_t runs = (uint64_t)p.n_bitems * p.span_stride;
        const uint64_t max_entries = (uint64_t)p.n_bitems * kBucketEntries;
        p.max_parts = (uint32_t)(p.n_slices + max_entries / kPart);
        p.run_cnt_off = 0;
        p.run_off_off = align(p.run_cnt_off + 4 * runs);
        p.slice_np_off = align(p.run_off_off + 4 * runs);
        p.meta_off = align(p.slice_np_off + 4 * (uint64_t)p.n_slices);
        p.parts_off = align(p.meta_off + 8);
        p.part_slice_off = align(p.parts_off + 4 * ((uint64_t)p.n_slices + 1));
        p.partials_off = align(p.part_slice_off + 4 * (uint64_t)p.max_parts);
        p.entries_off = align(p.partials_off + (uint64_t)p.max_parts * p.R * d.F * 4);
        p.bytes = p.entries_off + max_entries * 4 * ew;
        return p;
    }
    p.n_chunks = std::max<uint32_t>(1, div_up(d.n_points, kChunk));
    p.n_items = d.L * p.n_chunks;
    p.M = (uint64_t)p.n_slices * p.n_items;
    p.scan_blocks = (p.M + kScanThreads * kScanPer - 1) / (kScanThreads * kScanPer);
    const uint64_t max_entries = (uint64_t)d.n_points * d.L * 8;
    // a slice with c entries takes max(1, ceil(c / kPart)) parts
    p.max_parts = (uint32_t)(p.n_slices + max_entries / kPart);
    p.counts_off = 0;
    p.sums_off = align(p.counts_off + 4 * p.M);
    p.meta_off = align(p.sums_off + 4 * p.scan_blocks);        // [0] total entries, [1] total parts
    p.parts_off = align(p.meta_off + 8);                          // first part of each slice
    p.part_slice_off = align(p.parts_off + 4 * ((uint64_t)p.n_slices + 1));  // slice of each part
    p.partials_off = align(p.part_slice_off + 4 * (uint64_t)p.max_parts);
    p.entries_off = align(p.partials_off + (uint64_t)p.max_parts * p.R * d.F * 4);
    p.bytes = p.entries_off + max_entries * 4 * ew;
    return p;
}

// Lane shifts on the vector ALU (DPP), no LDS: lanes without a source, or in rows
// the row mask leaves out, get `fill`.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138, kRowBcast15 = 0x142, kRowBcast31 = 0x143;

template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v, uint32_t fill) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, kCtrl, kRowMask, 0xf, false);
}

// One step of the segmented inclusive sum: combine with the pair the shift brings
// (0 and flag 0 where there is none, which changes nothing).
template <int kCtrl, int kRowMask, int F>
__device__ __forceinline__ void seg_step(float *v, uint32_t &fl) {
    const uint32_t fin = dpp_u32<kCtrl, kRowMask>(fl, 0u);
#pragma unroll
    for (int f = 0; f < F; ++f) {
        const float vin = __uint_as_float(dpp_u32<kCtrl, kRowMask>(__float_as_uint(v[f]), 0u));
        v[f] = fl ? v[f] : v[f] + vin;
    }
    fl |= fin;
}

// The run-merged contributions of one point per lane at level l: the wavefront's
// 64 lanes hold 64 consecutive points; for each corner, lanes whose row equals the
// previous lane's join its run, and emit(row, v) is called on each run's last lane
// with the run's summed values.  Inactive lanes (past the last point) get keys no
// active lane has, so they form no runs.
template <int F, bool kValues, typename Emit>
__device__ __forceinline__ void point_runs(const HashGridDescriptor &d, uint32_t l, float scale, uint32_t res,
                                           bool dense, bool pow2, const float *xyz, const float *gin, bool active,
                                           Emit emit) {
    const int lane = lane_id();
    const Divisor two_b = make_divisor(2.0f * d.bound);
    const float px = divide(two_b, xyz[0] + d.bound);
    const float py = divide(two_b, xyz[1] + d.bound);
    const float pz = divide(two_b, xyz[2] + d.bound);
    uint32_t idx[8];
    float w[8];
    level_vertices(px, py, pz, scale, res, dense, d.T, pow2, d.offsets[l], d.n_rows - 1u, idx, w);
    float g[F];
#pragma unroll
    for (int f = 0; f < F; ++f) g[f] = kValues && active ? gin[f] : 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t key = active ? idx[c] : 0xffffffffu - (uint32_t)lane;
        float v[F];
#pragma unroll
        for (int f = 0; f < F; ++f) v[f] = w[c] * g[f];
        const uint32_t prev = dpp_u32<kWaveShr1, 0xf>(key, ~key);  // lane 0: no predecessor
        const bool head = prev != key;
        bool tail = true;
        if (__ballot(!head) != 0) {
            if (kValues) {
                // segmented inclusive sums of runs: (v, flag) pairs combined by
                // (a, fa) + (b, fb) = (fb ? b : a + b, fa | fb), shifts within
                // 16-lane rows, then row 15 -> rows 1 and 3, lane 31 -> rows 2 and 3
                uint32_t fl = head ? 1u : 0u;
                seg_step<kRowShr1, 0xf, F>(v, fl);
                seg_step<kRowShr2, 0xf, F>(v, fl);
                seg_step<kRowShr4, 0xf, F>(v, fl);
                seg_step<kRowShr8, 0xf, F>(v, fl);
                seg_step<kRowBcast15, 0xa, F>(v, fl);
                seg_step<kRowBcast31, 0xc, F>(v, fl);
            }
            const uint32_t next_head = dpp_u32<kWaveShl1, 0xf>(head ? 1u : 0u, 1u);  // lane 63: end of wave
            tail = next_head != 0u;
        }
        if (tail && active) emit(key, v, c);
    }
}

// Visits the run-merged contributions of work item `item` (level l, kChunk
// consecutive points) and calls emit(row, v, corner) on each run's last lane.
// The same inputs always give the same multiset of emitted rows, so counting and
// scattering agree.
template <int F, bool kValues, typename Emit>
__device__ __forceinline__ void item_contributions(const HashGridDescriptor &d, const BwdPlan &p, uint32_t item,
                                                   const float *__restrict__ pos, const float *__restrict__ dL_denc,
                                                   Emit emit) {
    const uint32_t l = item / p.n_chunks, k = item % p.n_chunks;
    const bool pow2 = (d.T & (d.T - 1u)) == 0;
    const bool level_major = d.enc_layout == NGP_ENC_LEVEL_MAJOR;
    const float scale = d.scales[l];
    const uint32_t res = d.resolutions[l];
    const bool dense = l < d.first_hash_level;
    constexpr int kSubs = kChunk / kItemThreads;
    // the inputs of a group of sub-chunks are all loaded before any is used, so
    // the loads overlap; 8 features per level hold half as many in flight (all
    // eight sub-chunks of them would not fit the registers)
    constexpr int kGroup = F >= 8 ? kSubs / 2 : kSubs;
    for (int g0 = 0; g0 < kSubs; g0 += kGroup) {
        float xyz[kGroup][3], gin[kGroup][F];
#pragma unroll
        for (int sub = 0; sub < kGroup; ++sub) {
            const uint32_t n = k * kChunk + (g0 + sub) * kItemThreads + threadIdx.x;
            const uint32_t np = n < d.n_points ? n : d.n_points - 1;
#pragma unroll
            for (int a = 0; a < 3; ++a) xyz[sub][a] = pos[3 * (size_t)np + a];
            if (kValues) {
                const size_t gb =
                    level_major ? ((size_t)l * d.n_points + np) * F : (size_t)np * d.L * F + (size_t)l * F;
#pragma unroll
                for (int f = 0; f < F; ++f) gin[sub][f] = dL_denc[gb + f];
            }
        }
#pragma unroll
        for (int sub = 0; sub < kGroup; ++sub) {
            const uint32_t n = k * kChunk + (g0 + sub) * kItemThreads + threadIdx.x;
            point_runs<F, kValues>(d, l, scale, res, dense, pow2, xyz[sub], gin[sub], n < d.n_points, emit);
        }
    }
}

// 1. per item, entries per slice (slice-major counters)
template <int F>
__global__ void __launch_bounds__(kItemThreads)
    bwd_count_kernel(HashGridDescriptor d, BwdPlan p, const float *__restrict__ pos, uint32_t *__restrict__ counts) {
    constexpr uint32_t R = slice_rows<F>();
    __shared__ uint32_t hist[kMaxSlices];
    for (uint32_t s = threadIdx.x; s < p.n_slices; s += kItemThreads) hist[s] = 0u;
    __syncthreads();
    const uint32_t item = blockIdx.x;
    item_contributions<F, false>(d, p, item, pos, nullptr,
                                 [&](uint32_t row, const float *, int) { atomicAdd(&hist[row / R], 1u); });
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < p.n_slices; s += kItemThreads) counts[(size_t)s * p.n_items + item] = hist[s];
}

// 2a. per-block sums of kScanThreads * kScanPer counters
__global__ void __launch_bounds__(kScanThreads) scan_sums_kernel(uint64_t M, const uint32_t *__restrict__ a,
                                                                 uint32_t *__restrict__ sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanThreads * kScanPer;
    uint32_t v = 0;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        const uint64_t i = base + (uint64_t)q * kScanThreads + threadIdx.x;
        if (i < M) v += a[i];
    }
    __shared__ uint32_t ws[kScanThreads / kWave];
    v = wave_sum(v);
    if (lane_id() == 0) ws[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

// 2b. exclusive scan inside each block plus the block's prefix (the sum of the
// block sums before it, added by the workgroup itself), in place; the last block
// writes the total entries into meta[0]
__global__ void __launch_bounds__(kScanThreads) scan_apply_kernel(uint64_t M, uint32_t *__restrict__ a,
                                                                  const uint32_t *__restrict__ sums,
                                                                  uint32_t *__restrict__ meta) {
    __shared__ uint32_t ws[kScanThreads / kWave];
    __shared__ uint32_t prefix;
    uint32_t v = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kScanThreads) v += sums[j];
    v = wave_sum(v);
    if (lane_id() == 0) ws[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) t += ws[w];
        prefix = t;
    }
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kScanThreads * kScanPer;
    uint32_t carry = prefix;
    for (int q = 0; q < kScanPer; ++q) {
        const uint64_t i = base + (uint64_t)q * kScanThreads + threadIdx.x;
        const uint32_t x = i < M ? a[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<kScanThreads>(x, &tot);
        if (i < M) a[i] = carry + ex;
        carry += tot;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) meta[0] = carry;
}

__device__ __forceinline__ uint32_t slice_begin(const BwdPlan &p, const uint32_t *offsets, const uint32_t *meta,
                                                uint32_t s) {
    return s < p.n_slices ? offsets[(size_t)s * p.n_items] : meta[0];
}

// 2d. one workgroup: parts per slice (kPart entries each, at least one) and their
// prefix, and the slice of every part; parts[n_slices] and meta[1] hold the total.
__global__ void __launch_bounds__(kScanThreads) plan_parts_kernel(BwdPlan p, const uint32_t *__restrict__ offsets,
                                                                  uint32_t *__restrict__ meta,
                                                                  uint32_t *__restrict__ parts,
                                                                  uint32_t *__restrict__ part_slice) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < p.n_slices; c += kScanThreads) {
        const uint32_t s = c + threadIdx.x;
        uint32_t k = 0;
        if (s < p.n_slices) {
            const uint32_t cnt = slice_begin(p, offsets, meta, s + 1) - slice_begin(p, offsets, meta, s);
            k = cnt == 0 ? 1u : (cnt + kPart - 1) / kPart;
        }
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<kScanThreads>(k, &tot);
        if (s < p.n_slices) {
            parts[s] = carry + ex;
            for (uint32_t j = 0; j < k; ++j) part_slice[carry + ex + j] = s;
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        parts[p.n_slices] = carry;
        meta[1] = carry;
    }
}

// 3. entries at each item's offsets
template <int F, bool kHalf>
__global__ void __launch_bounds__(kItemThreads)
    bwd_scatter_kernel(HashGridDescriptor d, BwdPlan p, const float *__restrict__ pos, const float *__restrict__ dL_denc,
                       const uint32_t *__restrict__ offsets, uint32_t *__restrict__ entries) {
    constexpr int EW = entry_words<F, kHalf>();
    constexpr uint32_t R = slice_rows<F>();
    __shared__ uint32_t cursor[kMaxSlices];
    const uint32_t item = blockIdx.x;
    for (uint32_t s = threadIdx.x; s < p.n_slices; s += kItemThreads) cursor[s] = offsets[(size_t)s * p.n_items + item];
    __syncthreads();
    item_contributions<F, true>(d, p, item, pos, dL_denc, [&](uint32_t row, const float *v, int) {
        const uint32_t s = row / R;
        const uint32_t at = atomicAdd(&cursor[s], 1u);
        uint32_t w[EW];
#pragma unroll
        for (int k = 0; k < EW; ++k) w[k] = 0u;
        w[0] = row - s * R;
        pack_values<F, kHalf>(w, v);
        uint4 *e = reinterpret_cast<uint4 *>(entries + (size_t)at * EW);
        if constexpr (EW == 2) {
            *reinterpret_cast<uint2 *>(entries + (size_t)at * EW) = make_uint2(w[0], w[1]);
        } else if constexpr (EW == 3) {
            *reinterpret_cast<uint3 *>(entries + (size_t)at * EW) = make_uint3(w[0], w[1], w[2]);
        } else {
#pragma unroll
            for (int q = 0; q < EW / 4; ++q) e[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
        }
    });
}


// One-pass table gradient, part 1 (F <= 2).  Block = item (level l, kBucketPoints
// consecutive points, one per thread and phase).  The item's run-merged
// contributions are computed once and kept in registers; their count per slice of
// the level goes to LDS, an exclusive scan of the counts places each slice's run
// in the item's own region of the entry buffer (kBucketEntries entries), and the
// run's start and length are stored for the reducers.  Within a run, entries are in
// LDS-atomic order; the reducer's sums do not depend on it beyond f64 rounding.
template <int F, bool kHalf>
__global__ void __launch_bounds__(kBucketThreads)
    bwd_bucket_kernel(HashGridDescriptor d, BwdPlan p, const float *__restrict__ pos,
                      const float *__restrict__ dL_denc, uint32_t *__restrict__ run_cnt,
                      uint32_t *__restrict__ run_off, uint32_t *__restrict__ entries) {
    constexpr int EW = entry_words<F, kHalf>();
    constexpr uint32_t R = slice_rows<F>();
    static_assert(kStageSlices == 2 * kBucketThreads, "two slice counters per thread");
    __shared__ uint32_t cnt[kStageSlices];
    __shared__ uint32_t cur[kStageSlices];
    const uint32_t item = blockIdx.x;
    const uint32_t l = item / p.n_bchunks, k = item % p.n_bchunks;
    const uint32_t s_lo = p.s_lo[l];
    const uint32_t span = p.s_hi[l] - s_lo + 1u;  // <= kStageSlices (host check)
    cnt[2 * threadIdx.x] = 0u;
    cnt[2 * threadIdx.x + 1] = 0u;
    const bool pow2 = (d.T & (d.T - 1u)) == 0;
    const bool level_major = d.enc_layout == NGP_ENC_LEVEL_MAJOR;
    const float scale = d.scales[l];
    const uint32_t res = d.resolutions[l];
    const bool dense = l < d.first_hash_level;
    // both phases' inputs are loaded up front, so only the first phase waits on memory
    float xyz_all[kBucketPhases][3], gin_all[kBucketPhases][F];
#pragma unroll
    for (int ph = 0; ph < kBucketPhases; ++ph) {
        const uint32_t n = k * kBucketPoints + ph * kBucketThreads + threadIdx.x;
        const uint32_t np = n < d.n_points ? n : d.n_points - 1;
#pragma unroll
        for (int q = 0; q < 3; ++q) xyz_all[ph][q] = pos[3 * (size_t)np + q];
        const size_t gb = level_major ? ((size_t)l * d.n_points + np) * F : (size_t)np * d.L * F + (size_t)l * F;
#pragma unroll
        for (int f = 0; f < F; ++f) gin_all[ph][f] = dL_denc[gb + f];
    }
    __syncthreads();  // counters cleared
    uint32_t rows[kBucketPhases][8];
    float vals[kBucketPhases][8][F];
    uint32_t mask[kBucketPhases];
#pragma unroll
    for (int ph = 0; ph < kBucketPhases; ++ph) {
        const uint32_t n = k * kBucketPoints + ph * kBucketThreads + threadIdx.x;
        uint32_t m = 0;
        point_runs<F, true>(d, l, scale, res, dense, pow2, xyz_all[ph], gin_all[ph], n < d.n_points,
                            [&](uint32_t row, const float *v, int c) {
                                rows[ph][c] = row;
#pragma unroll
                                for (int f = 0; f < F; ++f) vals[ph][c][f] = v[f];
                                m |= 1u << c;
                            });
        mask[ph] = m;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if ((m >> c) & 1u) atomicAdd(&cnt[rows[ph][c] / R - s_lo], 1u);
    }
    __syncthreads();
    const uint32_t c0 = cnt[2 * threadIdx.x], c1 = cnt[2 * threadIdx.x + 1];
    uint32_t total;
    const uint32_t ex = block_exclusive_scan<kBucketThreads>(c0 + c1, &total);
    cur[2 * threadIdx.x] = ex;
    cur[2 * threadIdx.x + 1] = ex + c0;
    const size_t rb = (size_t)item * p.span_stride;
    if (2 * threadIdx.x < span) {
        run_cnt[rb + 2 * threadIdx.x] = c0;
        run_off[rb + 2 * threadIdx.x] = ex;
    }
    if (2 * threadIdx.x + 1 < span) {
        run_cnt[rb + 2 * threadIdx.x + 1] = c1;
        run_off[rb + 2 * threadIdx.x + 1] = ex + c0;
    }
    __syncthreads();
    uint32_t *region = entries + (size_t)item * kBucketEntries * EW;
#pragma unroll
    for (int ph = 0; ph < kBucketPhases; ++ph) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if ((mask[ph] >> c) & 1u) {
                const uint32_t sl = rows[ph][c] / R;
                const uint32_t at = atomicAdd(&cur[sl - s_lo], 1u);
                uint32_t w[EW];
#pragma unroll
                for (int q = 0; q < EW; ++q) w[q] = 0u;
                w[0] = rows[ph][c] - sl * R;
                pack_values<F, kHalf>(w, vals[ph][c]);
                uint32_t *dst = region + (size_t)at * EW;
                if constexpr (EW == 2) {
                    *reinterpret_cast<uint2 *>(dst) = make_uint2(w[0], w[1]);
                } else {
                    static_assert(EW == 3, "one-pass entries are 2 or 3 words");
                    *reinterpret_cast<uint3 *>(dst) = make_uint3(w[0], w[1], w[2]);
                }
            }
        }
    }
}

// One-pass table gradient, part 2: parts of slice s (kPart entries each, at least
// one), from the run lengths of every item whose level reaches s.
__global__ void __launch_bounds__(256)
    bwd_slice_parts_kernel(HashGridDescriptor d, BwdPlan p, const uint32_t *__restrict__ run_cnt,
                           uint32_t *__restrict__ slice_np) {
    const uint32_t s = blockIdx.x;
    uint32_t tot = 0;
    for (uint32_t l = 0; l < d.L; ++l) {
        if (s < p.s_lo[l] || s > p.s_hi[l]) continue;
        const uint32_t sl = s - p.s_lo[l];
        for (uint32_t k = threadIdx.x; k < p.n_bchunks; k += 256)
            tot += run_cnt[(size_t)(l * p.n_bchunks + k) * p.span_stride + sl];
    }
    __shared__ uint32_t ws[256 / kWave];
    tot = wave_sum(tot);
    if (lane_id() == 0) ws[threadIdx.x / kWave] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < 256 / kWave; ++w) t += ws[w];
        slice_np[s] = t == 0 ? 1u : (t + kPart - 1) / kPart;
    }
}

// One-pass table gradient, part 3 (one workgroup): prefix of the parts per slice and
// the slice of every part; parts[n_slices] and meta[1] hold the total.
__global__ void __launch_bounds__(kScanThreads) bwd_bucket_plan_kernel(BwdPlan p, const uint32_t *__restrict__ slice_np,
                                                                       uint32_t *__restrict__ meta,
                                                                       uint32_t *__restrict__ parts,
                                                                       uint32_t *__restrict__ part_slice) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < p.n_slices; c += kScanThreads) {
        const uint32_t s = c + threadIdx.x;
        const uint32_t k = s < p.n_slices ? slice_np[s] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<kScanThreads>(k, &tot);
        if (s < p.n_slices) {
            parts[s] = carry + ex;
            for (uint32_t j = 0; j < k; ++j) part_slice[carry + ex + j] = s;
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        parts[p.n_slices] = carry;
        meta[1] = carry;
    }
}

template <int EW>
__device__ __forceinline__ void load_entry(const uint32_t *__restrict__ src, uint32_t *w) {
    if constexpr (EW == 2) {
        const uint2 q = *reinterpret_cast<const uint2 *>(src);
        w[0] = q.x;
        w[1] = q.y;
    } else if constexpr (EW == 3) {
        const uint3 q = *reinterpret_cast<const uint3 *>(src);
        w[0] = q.x;
        w[1] = q.y;
        w[2] = q.z;
    } else {
#pragma unroll
        for (int q = 0; q < EW / 4; ++q) {
            const uint4 u = reinterpret_cast<const uint4 *>(src)[q];
            w[4 * q] = u.x;
            w[4 * q + 1] = u.y;
            w[4 * q + 2] = u.z;
            w[4 * q + 3] = u.w;
        }
    }
}

// 4. workgroup b sums part b's entries in LDS.  A slice of one part writes its
// rows of the gradient; a slice of several writes partial rows, summed in order
// by combine_kernel.  The LDS sums are f64 (each row rounded to f32 once at the
// end): f64 LDS atomics run several times faster than f32 ones here, and they
// keep full relative precision whatever the magnitudes, with inf and NaN
// propagating as in a float sum.
template <int F, bool kHalf>
__global__ void __launch_bounds__(kReduceThreads)
    bwd_reduce_kernel(HashGridDescriptor d, BwdPlan p, const uint32_t *__restrict__ offsets,
                      const uint32_t *__restrict__ meta, const uint32_t *__restrict__ parts,
                      const uint32_t *__restrict__ part_slice, const uint32_t *__restrict__ entries,
                      float *__restrict__ partials, float *__restrict__ dtable) {
    constexpr int EW = entry_words<F, kHalf>();
    extern __shared__ double acc[];
    const uint32_t b = blockIdx.x;
    if (b >= meta[1]) return;
    const uint32_t s = part_slice[b];
    const uint32_t part = b - parts[s];
    const uint32_t n_parts = parts[s + 1] - parts[s];
    const uint32_t s_begin = slice_begin(p, offsets, meta, s), s_end = slice_begin(p, offsets, meta, s + 1);
    const uint32_t begin = s_begin + part * kPart;
    const uint32_t end = min(s_end, begin + kPart);
    const uint32_t rows = min(d.n_rows - s * p.R, p.R);
    // feature f of local row r at acc[f * R + r]: one atomic instruction's rows
    // spread over all banks
    for (uint32_t t = threadIdx.x; t < p.R * F; t += kReduceThreads) acc[t] = 0.0;
    __syncthreads();
    // kU entries in flight per thread: the loads are unconditional (indices past the
    // part clamp to its last entry and are dropped afterwards), so they issue back
    // to back instead of each waiting in its own branch
    constexpr int kU = 8;
    for (uint32_t e0 = begin; e0 < end; e0 += kU * kReduceThreads) {
        uint32_t w[kU][EW];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t e = min(e0 + u * kReduceThreads + threadIdx.x, end - 1u);
            load_entry<EW>(entries + (size_t)e * EW, w[u]);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool in = e0 + u * kReduceThreads + threadIdx.x < end;
#pragma unroll
            for (int f = 0; f < F; ++f) {
                const float v = entry_value<F, kHalf>(w[u], f);
                if (in && v != 0.0f) atomicAdd(&acc[f * p.R + w[u][0]], (double)v);
            }
        }
    }
    __syncthreads();
    float *out = n_parts == 1 ? dtable + (size_t)s * p.R * F : partials + (size_t)b * p.R * F;
    for (uint32_t t = threadIdx.x; t < rows * F; t += kReduceThreads) out[t] = (float)acc[(t % F) * p.R + t / F];
}

// One-pass table gradient, part 4: workgroup b sums part b of slice s, as
// bwd_reduce_kernel does, from the runs of the items whose level reaches s.  The
// items are visited in (level, chunk) order and a part takes those whose exclusive
// prefix of run lengths falls in [part * kPart, (part + 1) * kPart).  A wavefront
// reads kRunsPerWave runs side by side, two loads per lane and run in flight (a
// hashed level's run holds about 128 entries; longer runs loop), so eight entry
// loads per lane are in flight as in bwd_reduce_kernel.
// With kAdam, a slice summed in one part is not stored: its rows take the optimizer
// step in place (params, m, v), and bwd_combine does the same for the others.
template <int F, bool kHalf, bool kAdam = false>
__global__ void __launch_bounds__(kReduceThreads)
    bwd_reduce_runs_kernel(HashGridDescriptor d, BwdPlan p, const uint32_t *__restrict__ meta,
                           const uint32_t *__restrict__ parts, const uint32_t *__restrict__ part_slice,
                           const uint32_t *__restrict__ run_cnt, const uint32_t *__restrict__ run_off,
                           const uint32_t *__restrict__ entries, float *__restrict__ partials,
                           float *__restrict__ dtable, AdamStep adam = {}, float *__restrict__ adam_p = nullptr,
                           float *__restrict__ adam_m = nullptr, float *__restrict__ adam_v = nullptr) {
    constexpr int EW = entry_words<F, kHalf>();
    constexpr uint32_t R = slice_rows<F>();
    constexpr uint32_t kWaves = kReduceThreads / kWave;
    constexpr int kRunsPerWave = 4, kPerRun = 2;
    extern __shared__ double acc[];
    // selected runs: entry base (item region + run offset) and length
    __shared__ uint64_t sel_base[kReduceThreads];
    __shared__ uint32_t sel_cnt[kReduceThreads];
    const uint32_t b = blockIdx.x;
    if (b >= meta[1]) return;
    const uint32_t s = part_slice[b];
    const uint32_t part = b - parts[s];
    const uint32_t n_parts = parts[s + 1] - parts[s];
    const uint32_t rows = min(d.n_rows - s * R, R);
    for (uint32_t t = threadIdx.x; t < R * F; t += kReduceThreads) acc[t] = 0.0;
    __syncthreads();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    uint32_t carry = 0;  // entries of the items visited before
    for (uint32_t l = 0; l < d.L; ++l) {
        if (s < p.s_lo[l] || s > p.s_hi[l]) continue;
        const uint32_t sl = s - p.s_lo[l];
        for (uint32_t k0 = 0; k0 < p.n_bchunks; k0 += kReduceThreads) {
            const uint32_t k = k0 + threadIdx.x;
            const uint32_t item = l * p.n_bchunks + k;
            uint32_t c = 0, off = 0;
            if (k < p.n_bchunks) {
                c = run_cnt[(size_t)item * p.span_stride + sl];
                off = run_off[(size_t)item * p.span_stride + sl];
            }
            uint32_t tot, n_sel;
            const uint32_t ex = block_exclusive_scan<kReduceThreads>(c, &tot);
            const bool sel = c > 0 && (carry + ex) / kPart == part;
            const uint32_t at = block_exclusive_scan<kReduceThreads>(sel ? 1u : 0u, &n_sel);
            if (sel) {
                sel_base[at] = (uint64_t)item * kBucketEntries + off;
                sel_cnt[at] = c;
            }
            carry += tot;
            __syncthreads();
            for (uint32_t j0 = wave; j0 < n_sel; j0 += kWaves * kRunsPerWave) {
                uint64_t base[kRunsPerWave];
                uint32_t cnt[kRunsPerWave], longest = 0;
#pragma unroll
                for (int q = 0; q < kRunsPerWave; ++q) {
                    const uint32_t j = j0 + q * kWaves;
                    cnt[q] = j < n_sel ? sel_cnt[j] : 0u;
                    base[q] = j < n_sel ? sel_base[j] : 0ull;
                    longest = max(longest, cnt[q]);
                }
                for (uint32_t e0 = 0; e0 < longest; e0 += kPerRun * kWave) {
                    uint32_t w[kRunsPerWave][kPerRun][EW];
#pragma unroll
                    for (int q = 0; q < kRunsPerWave; ++q)
#pragma unroll
                        for (int h = 0; h < kPerRun; ++h) {
                            const uint32_t e = cnt[q] ? min(e0 + h * kWave + lane, cnt[q] - 1u) : 0u;
                            load_entry<EW>(entries + (base[q] + e) * EW, w[q][h]);
                        }
#pragma unroll
                    for (int q = 0; q < kRunsPerWave; ++q)
#pragma unroll
                        for (int h = 0; h < kPerRun; ++h) {
                            const bool in = e0 + h * kWave + lane < cnt[q];
#pragma unroll
                            for (int f = 0; f < F; ++f) {
                                const float v = entry_value<F, kHalf>(w[q][h], f);
                                if (in && v != 0.0f) atomicAdd(&acc[f * R + w[q][h][0]], (double)v);
                            }
                        }
                }
            }
            __syncthreads();  // the selection is rewritten by the next tile
        }
    }
    if (kAdam && n_parts == 1) {
        const size_t o = (size_t)s * R * F;
        // unrolled so that several rows' params, m and v loads are in flight at once
#pragma unroll 8
        for (uint32_t t = threadIdx.x; t < rows * F; t += kReduceThreads) {
            float mk = adam_m[o + t], vk = adam_v[o + t];
            adam_p[o + t] = adam_step(adam, adam_p[o + t], (float)acc[(t % F) * R + t / F], mk, vk);
            adam_m[o + t] = mk;
            adam_v[o + t] = vk;
        }
        return;
    }
    float *out = n_parts == 1 ? dtable + (size_t)s * R * F : partials + (size_t)b * R * F;
    for (uint32_t t = threadIdx.x; t < rows * F; t += kReduceThreads) out[t] = (float)acc[(t % F) * R + t / F];
}

// 5. slices of several parts: rows = sum of the parts' partial rows, in part
// order.  Block (s, c) covers 1/kCombineSplit of slice s.
constexpr uint32_t kCombineSplit = 16;
// With kAdam the sums are not stored: the rows take the optimizer step in place.
template <int F, bool kAdam = false>
__global__ void __launch_bounds__(256)
    bwd_combine_kernel(HashGridDescriptor d, BwdPlan p, const uint32_t *__restrict__ parts,
                       const float *__restrict__ partials, float *__restrict__ dtable, AdamStep adam = {},
                       float *__restrict__ adam_p = nullptr, float *__restrict__ adam_m = nullptr,
                       float *__restrict__ adam_v = nullptr) {
    const uint32_t s = blockIdx.x;
    const uint32_t p0 = parts[s], n_parts = parts[s + 1] - p0;
    if (n_parts <= 1) return;
    const uint32_t n = min(d.n_rows - s * p.R, p.R) * F;  // floats of the slice
    const uint32_t per = (p.R * F) / kCombineSplit;      // a multiple of 4
    const uint32_t lo = blockIdx.y * per, hi = min(n, lo + per);
    const size_t pstride = (size_t)p.R * F;
    const float *src = partials + (size_t)p0 * pstride;
    const size_t o = (size_t)s * p.R * F;
    for (uint32_t t = lo + 4 * threadIdx.x; t < hi; t += 4 * 256) {
        if (t + 4 <= hi) {
            float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (uint32_t q = 0; q < n_parts; ++q) {
                const float4 v = *reinterpret_cast<const float4 *>(src + q * pstride + t);
                a.x += v.x;
                a.y += v.y;
                a.z += v.z;
                a.w += v.w;
            }
            if constexpr (kAdam) {
                float4 p4 = *reinterpret_cast<const float4 *>(adam_p + o + t);
                float4 m4 = *reinterpret_cast<const float4 *>(adam_m + o + t);
                float4 v4 = *reinterpret_cast<const float4 *>(adam_v + o + t);
                p4.x = adam_step(adam, p4.x, a.x, m4.x, v4.x);
                p4.y = adam_step(adam, p4.y, a.y, m4.y, v4.y);
                p4.z = adam_step(adam, p4.z, a.z, m4.z, v4.z);
                p4.w = adam_step(adam, p4.w, a.w, m4.w, v4.w);
                *reinterpret_cast<float4 *>(adam_m + o + t) = m4;
                *reinterpret_cast<float4 *>(adam_v + o + t) = v4;
                *reinterpret_cast<float4 *>(adam_p + o + t) = p4;
            } else {
                *reinterpret_cast<float4 *>(dtable + o + t) = a;
            }
        } else {
            for (uint32_t u = t; u < hi; ++u) {
                float a = 0.0f;
                for (uint32_t q = 0; q < n_parts; ++q) a += src[q * pstride + u];
                if constexpr (kAdam) {
                    float mk = adam_m[o + u], vk = adam_v[o + u];
                    adam_p[o + u] = adam_step(adam, adam_p[o + u], a, mk, vk);
                    adam_m[o + u] = mk;
                    adam_v[o + u] = vk;
                } else {
                    dtable[o + u] = a;
                }
            }
        }
    }
}

// The optimizer step of rows no point reached (a zero gradient), for the fused
// update with no points.
__global__ void __launch_bounds__(256)
    adam_zero_grad_kernel(uint64_t n, AdamStep adam, float *__restrict__ adam_p, float *__restrict__ adam_m,
                          float *__restrict__ adam_v) {
    for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (uint64_t)gridDim.x * 256) {
        float mk = adam_m[k], vk = adam_v[k];
        adam_p[k] = adam_step(adam, adam_p[k], 0.0f, mk, vk);
        adam_m[k] = mk;
        adam_v[k] = vk;
    }
}

// `fused` (F <= 2, f32 table, one-pass plan only): the optimizer step in place of
// the gradient's store; dtable is then not written.
struct FusedAdam {
    AdamStep step;
    float *p, *m, *v;
};

template <int F, bool kHalf>
bool launch_backward(const HashGridDescriptor &d, hipStream_t stream, const float *pos, const float *g, float *dtable,
                     uint8_t *ws, const char *target, const FusedAdam *fused = nullptr) {
    const BwdPlan p = make_plan(d);
    uint32_t *meta = reinterpret_cast<uint32_t *>(ws + p.meta_off);
    uint32_t *parts = reinterpret_cast<uint32_t *>(ws + p.parts_off);
    uint32_t *part_slice = reinterpret_cast<uint32_t *>(ws + p.part_slice_off);
    float *partials = reinterpret_cast<float *>(ws + p.partials_off);
    uint32_t *entries = reinterpret_cast<uint32_t *>(ws + p.entries_off);
    const size_t lds = (size_t)p.R * F * sizeof(double);
    if constexpr (F <= 2) {
        if (p.bucket) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&bwd_reduce_runs_kernel<F, kHalf>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) {
                set_error(kLaunchFailed, "%s: cannot reserve %zu bytes of LDS: %s", target, lds, hipGetErrorString(e));
                return false;
            }
            uint32_t *run_cnt = reinterpret_cast<uint32_t *>(ws + p.run_cnt_off);
            uint32_t *run_off = reinterpret_cast<uint32_t *>(ws + p.run_off_off);
            uint32_t *slice_np = reinterpret_cast<uint32_t *>(ws + p.slice_np_off);
            hipLaunchKernelGGL((bwd_bucket_kernel<F, kHalf>), dim3(p.n_bitems), dim3(kBucketThreads), 0, stream, d, p,
                               pos, g, run_cnt, run_off, entries);
            hipLaunchKernelGGL(bwd_slice_parts_kernel, dim3(p.n_slices), dim3(256), 0, stream, d, p, run_cnt, slice_np);
            hipLaunchKernelGGL(bwd_bucket_plan_kernel, dim3(1), dim3(kScanThreads), 0, stream, p, slice_np, meta, parts,
                               part_slice);
            if constexpr (!kHalf) {
                if (fused != nullptr) {
                    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&bwd_reduce_runs_kernel<F, false, true>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                    if (e != hipSuccess) {
                        set_error(kLaunchFailed, "%s: cannot reserve %zu bytes of LDS: %s", target, lds,
                                  hipGetErrorString(e));
                        return false;
                    }
                    hipLaunchKernelGGL((bwd_reduce_runs_kernel<F, false, true>), dim3(p.max_parts),
                                       dim3(kReduceThreads), lds, stream, d, p, meta, parts, part_slice, run_cnt,
                                       run_off, entries, partials, dtable, fused->step, fused->p, fused->m, fused->v);
                    hipLaunchKernelGGL((bwd_combine_kernel<F, true>), dim3(p.n_slices, kCombineSplit), dim3(256), 0,
                                       stream, d, p, parts, partials, dtable, fused->step, fused->p, fused->m,
                                       fused->v);
                    return true;
                }
            }
            hipLaunchKernelGGL((bwd_reduce_runs_kernel<F, kHalf>), dim3(p.max_parts), dim3(kReduceThreads), lds, stream,
                               d, p, meta, parts, part_slice, run_cnt, run_off, entries, partials, dtable);
            hipLaunchKernelGGL(bwd_combine_kernel<F>, dim3(p.n_slices, kCombineSplit), dim3(256), 0, stream, d, p,
                               parts, partials, dtable);
            return true;
        }
    }
    if (fused != nullptr) {
        set_error(kBadDescriptor, "%s: the fused update needs the one-pass table gradient", target);
        return false;
    }
    uint32_t *counts = reinterpret_cast<uint32_t *>(ws + p.counts_off);
    uint32_t *sums = reinterpret_cast<uint32_t *>(ws + p.sums_off);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&bwd_reduce_kernel<F, kHalf>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) {
        set_error(kLaunchFailed, "%s: cannot reserve %zu bytes of LDS: %s", target, lds, hipGetErrorString(e));
        return false;
    }
    hipLaunchKernelGGL(bwd_count_kernel<F>, dim3(p.n_items), dim3(kItemThreads), 0, stream, d, p, pos, counts);
    hipLaunchKernelGGL(scan_sums_kernel, dim3((uint32_t)p.scan_blocks), dim3(kScanThreads), 0, stream, p.M, counts,
                       sums);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((uint32_t)p.scan_blocks), dim3(kScanThreads), 0, stream, p.M, counts,
                       sums, meta);
    hipLaunchKernelGGL(plan_parts_kernel, dim3(1), dim3(kScanThreads), 0, stream, p, counts, meta, parts,
                       part_slice);
    hipLaunchKernelGGL((bwd_scatter_kernel<F, kHalf>), dim3(p.n_items), dim3(kItemThreads), 0, stream, d, p, pos, g,
                       counts, entries);
    hipLaunchKernelGGL((bwd_reduce_kernel<F, kHalf>), dim3(p.max_parts), dim3(kReduceThreads), lds, stream, d, p, counts,
                       meta, parts, part_slice, entries, partials, dtable);
    hipLaunchKernelGGL(bwd_combine_kernel<F>, dim3(p.n_slices, kCombineSplit), dim3(256), 0, stream, d, p, parts,
                       partials, dtable);
    return true;
}

__global__ void zero_table_kernel(float4 *__restrict__ a, size_t n4, float *__restrict__ tail, size_t n_tail) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t j = g; j < n4; j += stride) a[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (g < n_tail) tail[g] = 0.0f;
}

bool validate(const HashGridDescriptor &d, const char *target) {
    if (d.L == 0 || d.L > NGP_HASHGRID_MAX_LEVELS) {
        set_error(kBadDescriptor, "%s: L=%u must be in [1, %d]", target, d.L, NGP_HASHGRID_MAX_LEVELS);
        return false;
    }
    if (!(d.F == 1 || d.F == 2 || d.F == 4 || d.F == 8)) {
        set_error(kBadDescriptor, "%s: F=%u must be 1, 2, 4 or 8", target, d.F);
        return false;
    }
    if (d.T == 0 || d.first_hash_level > d.L || d.n_rows == 0 || d.offsets[0] != 0 || d.offsets[d.L] != d.n_rows ||
        !(d.bound > 0.0f)) {
        set_error(kBadDescriptor, "%s: inconsistent level table (T=%u first_hash_level=%u n_rows=%u bound=%g)", target,
                  d.T, d.first_hash_level, d.n_rows, (double)d.bound);
        return false;
    }
    for (uint32_t l = 0; l < d.L; ++l) {
        if (d.offsets[l + 1] < d.offsets[l]) {
            set_error(kBadDescriptor, "%s: offsets must be non-decreasing", target);
            return false;
        }
    }
    if (d.enc_layout != NGP_ENC_POINT_MAJOR && d.enc_layout != NGP_ENC_LEVEL_MAJOR) {
        set_error(kBadDescriptor, "%s: enc_layout=%u must be 0 (point-major) or 1 (level-major)", target, d.enc_layout);
        return false;
    }
    if (d.table_dtype != NGP_DTYPE_F32 && d.table_dtype != NGP_DTYPE_F16) {
        set_error(kBadDescriptor, "%s: table_dtype=%u must be 0 (f32) or 1 (f16)", target, d.table_dtype);
        return false;
    }
    return true;
}

template <typename TT>
void launch_fwd(const HashGridDescriptor &d, hipStream_t s, const float *pos, const void *table, float *enc) {
    const TT *t = static_cast<const TT *>(table);
    if (d.enc_layout == NGP_ENC_LEVEL_MAJOR) {
        const uint32_t g = level_grid(d.n_points, d.L);
        switch (d.F) {
            case 1: hipLaunchKernelGGL((hashgrid_fwd_lm_kernel<1, TT>), dim3(g), dim3(kBlock), 0, s, d, pos, t, enc); break;
            case 2: hipLaunchKernelGGL((hashgrid_fwd_lm_kernel<2, TT>), dim3(g), dim3(kBlock), 0, s, d, pos, t, enc); break;
            case 4: hipLaunchKernelGGL((hashgrid_fwd_lm_kernel<4, TT>), dim3(g), dim3(kBlock), 0, s, d, pos, t, enc); break;
            default: hipLaunchKernelGGL((hashgrid_fwd_lm_kernel<8, TT>), dim3(g), dim3(kBlock), 0, s, d, pos, t, enc); break;
        }
        return;
    }
    const uint32_t grid = div_up((uint64_t)d.n_points * d.L, kBlock);
    switch (d.F) {
        case 1: hipLaunchKernelGGL((hashgrid_fwd_kernel<1, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
        case 2: hipLaunchKernelGGL((hashgrid_fwd_kernel<2, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
        case 4: hipLaunchKernelGGL((hashgrid_fwd_kernel<4, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
        default: hipLaunchKernelGGL((hashgrid_fwd_kernel<8, TT>), dim3(grid), dim3(kBlock), 0, s, d, pos, t, enc); break;
    }
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void hashgrid_encode(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "hashgrid_encode";
    HashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!validate(d, target)) return;
    if (d.n_points == 0) return;
    if (!check_buffers(buffers, 3, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    const float *pos = (const float *)buffers[0];
    float *enc = (float *)buffers[2];
    if (d.table_dtype == NGP_DTYPE_F32) launch_fwd<float>(d, stream, pos, buffers[1], enc);
    else launch_fwd<__half>(d, stream, pos, buffers[1], enc);
    check_launch(target);
}

extern "C" int hashgrid_encode_backward_adam_supported(const char *opaque, size_t opaque_len) {
    const char *target = "hashgrid_encode_backward_adam_supported";
    HashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return 0;
    if (!validate(d, target)) return 0;
    return d.F <= 2 && d.table_dtype == NGP_DTYPE_F32 && make_plan(d).bucket ? 1 : 0;
}

extern "C" void hashgrid_encode_backward_adam(ngp_stream_t stream_, void **buffers, const char *opaque,
                                              size_t opaque_len) {
    const char *target = "hashgrid_encode_backward_adam";
    static_assert(offsetof(HashGridAdamDescriptor, adam) == sizeof(HashGridDescriptor), "no padding");
    HashGridAdamDescriptor q;
    if (!read_descriptor(opaque, opaque_len, &q, target)) return;
    const HashGridDescriptor &d = q.grid;
    if (!validate(d, target)) return;
    if (d.F > 2 || d.table_dtype != NGP_DTYPE_F32 || q.adam.flags != 0u || q.adam.mirror_dtype != NGP_DTYPE_F32 ||
        q.adam.n != (uint64_t)d.n_rows * d.F) {
        set_error(kBadDescriptor,
                  "%s: needs F <= 2, an f32 table, no loss scaling or f16 mirror, and adam.n = n_rows * F", target);
        return;
    }
    const BwdPlan plan = make_plan(d);
    if (!plan.bucket || plan.n_slices > kMaxSlices) {
        set_error(kBadDescriptor, "%s: the table's levels must fit the one-pass table gradient", target);
        return;
    }
    if ((uint64_t)d.n_points * d.L * 8 >= 0xffffffffull) {
        set_error(kBadDescriptor, "%s: n_points * L * 8 must stay below 2^32", target);
        return;
    }
    if (!check_buffer_list(buffers, {2, 3, 4}, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    FusedAdam f{make_adam_step(q.adam), (float *)buffers[2], (float *)buffers[3], (float *)buffers[4]};
    if (d.n_points == 0) {
        const uint64_t n = q.adam.n;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384));
        hipLaunchKernelGGL(adam_zero_grad_kernel, dim3(grid), dim3(256), 0, stream, n, f.step, f.p, f.m, f.v);
        check_launch(target);
        return;
    }
    if (!check_buffers(buffers, 6, target)) return;
    uint8_t *ws = (uint8_t *)buffers[5];
    if ((reinterpret_cast<uintptr_t>(ws) & 255u) != 0 ||
        ((reinterpret_cast<uintptr_t>(f.p) | reinterpret_cast<uintptr_t>(f.m) | reinterpret_cast<uintptr_t>(f.v)) &
         15u) != 0) {
        set_error(kBadArgument, "%s: params, m and v must be 16-byte and the workspace 256-byte aligned", target);
        return;
    }
    const float *pos = (const float *)buffers[0];
    const float *g = (const float *)buffers[1];
    const bool ok = d.F == 1 ? launch_backward<1, false>(d, stream, pos, g, nullptr, ws, target, &f)
                             : launch_backward<2, false>(d, stream, pos, g, nullptr, ws, target, &f);
    if (!ok) return;
    check_launch(target);
}

extern "C" size_t hashgrid_encode_backward_workspace_size(const char *opaque, size_t opaque_len) {
    HashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, "hashgrid_encode_backward_workspace_size")) return 0;
    if (!validate(d, "hashgrid_encode_backward_workspace_size")) return 0;
    return (size_t)make_plan(d).bytes;
}

extern "C" void hashgrid_encode_backward(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "hashgrid_encode_backward";
    HashGridDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!validate(d, target)) return;
    const BwdPlan plan = make_plan(d);
    if (plan.n_slices > kMaxSlices) {
        set_error(kBadDescriptor, "%s: table of %u rows exceeds %u row slices", target, d.n_rows, kMaxSlices);
        return;
    }
    if ((uint64_t)d.n_points * d.L * 8 >= 0xffffffffull) {
        set_error(kBadDescriptor, "%s: n_points * L * 8 must stay below 2^32", target);
        return;
    }
    if (!check_buffer_list(buffers, {2}, target)) return;
    if (d.n_points > 0 && !check_buffers(buffers, 4, target)) return;
    hipStream_t stream = (hipStream_t)stream_;
    const float *pos = (const float *)buffers[0];
    const float *g = (const float *)buffers[1];
    float *dtable = (float *)buffers[2];
    if (d.n_points == 0) {
        // no contributions: the gradient is zero
        const size_t n = (size_t)d.n_rows * d.F;
        const bool vec = (reinterpret_cast<uintptr_t>(dtable) & 15u) == 0;
        const size_t n4 = vec ? n / 4 : 0;
        const size_t n_tail = n - 4 * n4;
        const uint32_t zgrid =
            (uint32_t)std::max<size_t>(1, std::min<size_t>((std::max(n4, n_tail) + 255) / 256, 16384));
        hipLaunchKernelGGL(zero_table_kernel, dim3(zgrid), dim3(256), 0, stream, reinterpret_cast<float4 *>(dtable),
                           n4, dtable + 4 * n4, n_tail);
        check_launch(target);
        return;
    }
    uint8_t *ws = (uint8_t *)buffers[3];
    if ((reinterpret_cast<uintptr_t>(ws) & 255u) != 0 || (reinterpret_cast<uintptr_t>(dtable) & 15u) != 0) {
        set_error(kBadArgument, "%s: dL_dtable must be 16-byte and the workspace 256-byte aligned", target);
        return;
    }
    bool ok = false;
    const bool half = d.table_dtype == NGP_DTYPE_F16;
    switch (d.F) {
        case 1: ok = launch_backward<1, false>(d, stream, pos, g, dtable, ws, target); break;
        case 2:
            ok = half ? launch_backward<2, true>(d, stream, pos, g, dtable, ws, target)
                      : launch_backward<2, false>(d, stream, pos, g, dtable, ws, target);
            break;
        case 4:
            ok = half ? launch_backward<4, true>(d, stream, pos, g, dtable, ws, target)
                      : launch_backward<4, false>(d, stream, pos, g, dtable, ws, target);
            break;
        default: ok = launch_backward<8, false>(d, stream, pos, g, dtable, ws, target); break;
    }
    if (!ok) return;
    check_launch(target);
}
