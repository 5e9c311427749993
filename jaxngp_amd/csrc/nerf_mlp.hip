// nerf_mlp.hip — fused NeRF MLP (density MLP, SH(4) of the view direction,
// colour MLP, activations) on MFMA, forward and backward.
//
// Restates models/nerfs.py:39-86 with make_nerf_ngp's shapes (nerfs.py:422-454):
//   h1 = relu(x W1)            x: encodings [32], W1 [32,64]
//   o  = h1 W2                 W2 [64,16]; sigma = trunc_exp(o[0])
//   c  = [o, SH4(dir)]         SH4: encoders.py:384-405 (16 values)
//   h2 = relu(c W3)            W3 [32,64]
//   h3 = relu(h2 W4)           W4 [64,64]
//   rgb = sigmoid(h3 W5)       W5 [64,3]
// No biases (CoordinateBasedMLP, nerfs.py:89-128).  trunc_exp's VJP uses
// exp(clip(x, -15, 15)) (nerfs.py:221-238).
//
// Layout.  A wavefront owns a tile of 16 samples and works on transposed
// activations Z^T [features, 16 samples] with v_mfma_f32_16x16x4f32:
// lane (j, g) = (lane % 16, lane / 16) holds, in register r of block b, feature
// 16b + 4g + r of sample j — the MFMA output layout.  Fed back as the B operand
// of the next layer, K-slot g of K-step (b, r) then carries input feature
// 16b + 4g + r, so the A operand (weights) is read from LDS with that same
// permutation and no activation ever leaves registers between layers.
// Backward uses the same trick with W instead of W^T.  Weight gradients
// dW[in,out] = sum_s Z[s,in] dY[s,out] need samples on the K axis: the tile's Z and
// dY pass through a per-wave LDS transpose, and the dW blocks accumulate in
// registers across all tiles a wave visits; waves then add into LDS, each
// workgroup writes one row of the workspace, and a second kernel sums the rows
// in a fixed order.
#include <mutex>

#include "ngp_common.h"
#include "hashgrid_device.h"
#include "ngp_device.h"
#include "rng_device.h"

namespace ngp {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kThreads = kWaves * kWave;
constexpr int kTile = 16;

// flax parameter offsets (floats) in the flat parameter array
constexpr int P_W1 = 0, P_W2 = 2048, P_W3 = 3072, P_W4 = 5120, P_W5 = 9216, P_END = 9408;
static_assert(P_END == NGP_MLP_N_PARAMS, "parameter count");

// LDS copies: W[in][out_padded + 4].  A row stride of 4 mod 16 puts the four lane
// groups g of an MFMA operand read (rows 4g + r, columns i) 16 banks apart, so a
// forward layer's weight reads are free of bank conflicts (an odd stride put them
// 4 banks apart, four lanes on each bank).
constexpr int S_W1 = 68, S_W2 = 20, S_W3 = 68, S_W4 = 68, S_W5 = 20;
constexpr int L_W1 = 0;
constexpr int L_W2 = L_W1 + 32 * S_W1;
constexpr int L_W3 = L_W2 + 64 * S_W2;
constexpr int L_W4 = L_W3 + 32 * S_W3;
constexpr int L_W5 = L_W4 + 64 * S_W4;
constexpr int L_END = L_W5 + 64 * S_W5;  // 11264 floats
static_assert(S_W1 % 16 == 4 && S_W2 % 16 == 4 && S_W3 % 16 == 4 && S_W4 % 16 == 4 && S_W5 % 16 == 4,
              "weight row strides of 4 mod 16");

// per-wave transpose scratch of the backward pass: [16 samples][68]
constexpr int S_T = 68;
constexpr int T_FLOATS = kTile * S_T;

__device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ void stage_weights(float *lds, const float *__restrict__ params) {
    for (int t = threadIdx.x; t < L_END; t += kThreads) lds[t] = 0.0f;
    __syncthreads();
    for (int t = threadIdx.x; t < 32 * 64; t += kThreads) lds[L_W1 + (t / 64) * S_W1 + t % 64] = params[P_W1 + t];
    for (int t = threadIdx.x; t < 64 * 16; t += kThreads) lds[L_W2 + (t / 16) * S_W2 + t % 16] = params[P_W2 + t];
    for (int t = threadIdx.x; t < 32 * 64; t += kThreads) lds[L_W3 + (t / 64) * S_W3 + t % 64] = params[P_W3 + t];
    for (int t = threadIdx.x; t < 64 * 64; t += kThreads) lds[L_W4 + (t / 64) * S_W4 + t % 64] = params[P_W4 + t];
    for (int t = threadIdx.x; t < 64 * 3; t += kThreads) lds[L_W5 + (t / 3) * S_W5 + t % 3] = params[P_W5 + t];
    __syncthreads();
}

// Y^T = W^T Z^T.  `in` holds IN/16 blocks, `out` OUT/16 blocks.
template <int IN, int OUT, int STRIDE>
__device__ __forceinline__ void layer_fwd(const float *W, const v4f *in, v4f *out, int i, int g) {
#pragma unroll
    for (int ob = 0; ob < OUT / 16; ++ob) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < IN / 16; ++b) {
#pragma unroll
            for (int r = 0; r < 4; ++r) acc = mfma(W[(16 * b + 4 * g + r) * STRIDE + 16 * ob + i], in[b][r], acc);
        }
        out[ob] = acc;
    }
}

// dZ^T = W dY^T.
template <int IN, int OUT, int STRIDE>
__device__ __forceinline__ void layer_bwd(const float *W, const v4f *dout, v4f *din, int i, int g) {
#pragma unroll
    for (int ib = 0; ib < IN / 16; ++ib) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ob = 0; ob < OUT / 16; ++ob) {
#pragma unroll
            for (int r = 0; r < 4; ++r) acc = mfma(W[(16 * ib + i) * STRIDE + 16 * ob + 4 * g + r], dout[ob][r], acc);
        }
        din[ib] = acc;
    }
}

template <int N>
__device__ __forceinline__ void relu(v4f *z) {
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) z[b][r] = fmaxf(z[b][r], 0.0f);
}

// gradient through relu: keep where the forward output was positive
template <int N>
__device__ __forceinline__ void relu_mask(v4f *d, const v4f *z) {
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) d[b][r] = z[b][r] > 0.0f ? d[b][r] : 0.0f;
}

// The 4 SH(4) values 4g..4g+3 of unit direction (x, y, z) (encoders.py:384-405).
__device__ __forceinline__ v4f sh4_part(float x, float y, float z, int g) {
    const float xy = x * y, xz = x * z, yz = y * z, x2 = x * x, y2 = y * y, z2 = z * z;
    v4f v;
    if (g == 0) {
        v = {0.28209479177387814f, -0.48860251190291987f * y, 0.48860251190291987f * z, -0.48860251190291987f * x};
    } else if (g == 1) {
        v = {1.0925484305920792f * xy, -1.0925484305920792f * yz, 0.94617469575755997f * z2 - 0.31539156525251999f,
             -1.0925484305920792f * xz};
    } else if (g == 2) {
        v = {0.54627421529603959f * x2 - 0.54627421529603959f * y2, 0.59004358992664352f * y * (-3.0f * x2 + y2),
             2.8906114426405538f * xy * z, 0.45704579946446572f * y * (1.0f - 5.0f * z2)};
    } else {
        v = {0.3731763325901154f * z * (5.0f * z2 - 3.0f), 0.45704579946446572f * x * (1.0f - 5.0f * z2),
             1.4453057213202769f * z * (x2 - y2), 0.59004358992664352f * x * (-x2 + 3.0f * y2)};
    }
    return v;
}

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + det_expf(-x)); }

struct Fwd {
    v4f x[2], h1[4], c[2], h2[4], h3[4], rr;  // c[0] = o (density MLP output), c[1] = SH
};

// Features 16b + 4g .. 16b + 4g + 3 of sample s are levels 8b + 2g and 8b + 2g + 1
// (two features each).  Point-major: one 16-byte vector at enc[s*32 + 16b + 4g];
// level-major [16, n, 2]: two 8-byte pairs at enc[(level*n + s)*2].
__device__ __forceinline__ v4f load_enc4(const float *__restrict__ enc, bool level_major, uint32_t n, size_t s,
                                         int b, int g) {
    if (!level_major) return *reinterpret_cast<const v4f *>(enc + s * 32 + 16 * b + 4 * g);
    const size_t l0 = 8 * b + 2 * g;
    const float2 a = *reinterpret_cast<const float2 *>(enc + (l0 * n + s) * 2);
    const float2 c = *reinterpret_cast<const float2 *>(enc + ((l0 + 1) * n + s) * 2);
    return v4f{a.x, a.y, c.x, c.y};
}

__device__ __forceinline__ void store_enc4(float *__restrict__ enc, bool level_major, uint32_t n, size_t s, int b,
                                           int g, v4f v) {
    if (!level_major) {
        *reinterpret_cast<v4f *>(enc + s * 32 + 16 * b + 4 * g) = v;
        return;
    }
    const size_t l0 = 8 * b + 2 * g;
    *reinterpret_cast<float2 *>(enc + (l0 * n + s) * 2) = make_float2(v[0], v[1]);
    *reinterpret_cast<float2 *>(enc + ((l0 + 1) * n + s) * 2) = make_float2(v[2], v[3]);
}

// The inputs of one 16-sample tile: lane (i, g)'s encodings of sample s0 + i, its
// direction and (backward) its output gradient.  Loads are unconditional (samples
// past n read sample n - 1 and are then zeroed), so a tile's loads can be issued
// one tile ahead of its use.
struct TileIn {
    v4f x[2], gin;
    float dx, dy, dz;
};

__device__ __forceinline__ TileIn load_tile(const float *__restrict__ enc, bool level_major,
                                            const float *__restrict__ dirs, const float *__restrict__ grads,
                                            uint32_t n, size_t s0, int i, int g) {
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    TileIn t;
    if (n == 0) {  // nothing to load (the backward still runs to write zero rows)
        t.x[0] = t.x[1] = t.gin = zero;
        t.dx = t.dy = t.dz = 0.0f;
        return t;
    }
    const size_t s = s0 + i;
    const bool live = s < n;
    const size_t sc = live ? s : (size_t)n - 1;
#pragma unroll
    for (int b = 0; b < 2; ++b) t.x[b] = load_enc4(enc, level_major, n, sc, b, g);
    t.dx = dirs[3 * sc + 0];
    t.dy = dirs[3 * sc + 1];
    t.dz = dirs[3 * sc + 2];
    t.gin = grads != nullptr ? *reinterpret_cast<const v4f *>(grads + 4 * sc) : zero;
    if (!live) {
        t.x[0] = t.x[1] = t.gin = zero;
        t.dx = t.dy = t.dz = 0.0f;
    }
    return t;
}

// Forward of a tile whose encodings f.x are loaded, for direction (dx, dy, dz).
__device__ __forceinline__ void forward_from_x(const float *lds, float dx, float dy, float dz, int i, int g, Fwd &f) {
    layer_fwd<32, 64, S_W1>(lds + L_W1, f.x, f.h1, i, g);
    relu<4>(f.h1);
    layer_fwd<64, 16, S_W2>(lds + L_W2, f.h1, &f.c[0], i, g);
    f.c[1] = sh4_part(dx, dy, dz, g);
    layer_fwd<32, 64, S_W3>(lds + L_W3, f.c, f.h2, i, g);
    relu<4>(f.h2);
    layer_fwd<64, 64, S_W4>(lds + L_W4, f.h2, f.h3, i, g);
    relu<4>(f.h3);
    layer_fwd<64, 16, S_W5>(lds + L_W5, f.h3, &f.rr, i, g);
}


__global__ void __launch_bounds__(kThreads)
    mlp_forward_kernel(uint32_t n, bool level_major, const float *__restrict__ enc, const float *__restrict__ dirs,
                       const float *__restrict__ params, float *__restrict__ drgbs) {
    __shared__ float lds[L_END];
    stage_weights(lds, params);
    const int lane = lane_id();
    const int i = lane & 15, g = lane >> 4;
    const uint32_t n_tiles = (n + kTile - 1) / kTile;
    const uint32_t wave = blockIdx.x * kWaves + threadIdx.x / kWave;
    const uint32_t n_waves = gridDim.x * kWaves;
    TileIn next = load_tile(enc, level_major, dirs, nullptr, n, (size_t)min(wave, n_tiles - 1) * kTile, i, g);
    for (uint32_t t = wave; t < n_tiles; t += n_waves) {
        const size_t s0 = (size_t)t * kTile;
        const TileIn in = next;
        next = load_tile(enc, level_major, dirs, nullptr, n, (size_t)min(t + n_waves, n_tiles - 1) * kTile, i, g);
        Fwd f;
        f.x[0] = in.x[0];
        f.x[1] = in.x[1];
        forward_from_x(lds, in.dx, in.dy, in.dz, i, g, f);
        const size_t s = s0 + i;
        if (g == 0 && s < n) {
            const float sigma = det_expf(f.c[0][0]);
            v4f out = {sigma, sigmoidf(f.rr[0]), sigmoidf(f.rr[1]), sigmoidf(f.rr[2])};
            *reinterpret_cast<v4f *>(drgbs + 4 * s) = out;
        }
    }
}

// Writes a [16 samples x NB*16 features] tile held in MFMA output layout into a
// per-wave LDS scratch as [sample][feature].
template <int NB>
__device__ __forceinline__ void tile_to_lds(float *scr, const v4f *z, int i, int g) {
#pragma unroll
    for (int b = 0; b < NB; ++b) *reinterpret_cast<v4f *>(scr + i * S_T + 16 * b + 4 * g) = z[b];
}

// acc[ib*OB + ob] += Z^T dY over the 16 samples of the tile (both in scratch).
template <int IB, int OB>
__device__ __forceinline__ void dw_accumulate(v4f *acc, const float *zs, const float *ys, int i, int g) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float a[IB], bb[OB];
#pragma unroll
        for (int ib = 0; ib < IB; ++ib) a[ib] = zs[(4 * t + g) * S_T + 16 * ib + i];
#pragma unroll
        for (int ob = 0; ob < OB; ++ob) bb[ob] = ys[(4 * t + g) * S_T + 16 * ob + i];
#pragma unroll
        for (int ib = 0; ib < IB; ++ib)
#pragma unroll
            for (int ob = 0; ob < OB; ++ob) acc[ib * OB + ob] = mfma(a[ib], bb[ob], acc[ib * OB + ob]);
    }
}

// Adds dW blocks (MFMA layout: lane (j, g) reg r = dW[16ib + 4g + r][16ob + j]) into
// an LDS array laid out as the flax kernel [in][out_dim].  Every (row, col) is owned
// by exactly one lane of the calling wave.
template <int IB, int OB>
__device__ __forceinline__ void dw_to_lds(float *dst, const v4f *acc, int out_dim, int i, int g) {
#pragma unroll
    for (int ib = 0; ib < IB; ++ib)
#pragma unroll
        for (int ob = 0; ob < OB; ++ob) {
            const int col = 16 * ob + i;
            if (col < out_dim) {
#pragma unroll
                for (int r = 0; r < 4; ++r) dst[(16 * ib + 4 * g + r) * out_dim + col] += acc[ib * OB + ob][r];
            }
        }
}

// The workgroup's waves split each layer's dW: wave `part` owns column block
// ob = part of an (IB, 4)-block layer, or row block ib = part of an (4, 1)-block
// layer, and accumulates it over the tiles of all kWaves waves (their scratch
// areas, in wave order), so it holds a quarter of the weight gradient.
template <int IB, int OB>
__device__ __forceinline__ void dw_accumulate_part(v4f *acc, const float *scratch, int part, int i, int g) {
    static_assert(OB == 4 || (OB == 1 && IB == 4), "blocks split over four waves");
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const float *zs = scratch + w * 2 * T_FLOATS;
        const float *ys = zs + T_FLOATS;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int r0 = (4 * t + g) * S_T;
            if constexpr (OB == 4) {
                const float bb = ys[r0 + 16 * part + i];
#pragma unroll
                for (int ib = 0; ib < IB; ++ib) acc[ib] = mfma(zs[r0 + 16 * ib + i], bb, acc[ib]);
            } else {
                acc[0] = mfma(zs[r0 + 16 * part + i], ys[r0 + i], acc[0]);
            }
        }
    }
}

// Writes this wave's dW blocks into the flax-layout row [in][out_dim] (every
// element of the layer is owned by exactly one lane of one wave).
template <int IB, int OB>
__device__ __forceinline__ void dw_part_to_row(float *dst, const v4f *acc, int out_dim, int part, int i, int g) {
    if constexpr (OB == 4) {
#pragma unroll
        for (int ib = 0; ib < IB; ++ib)
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[(16 * ib + 4 * g + r) * out_dim + 16 * part + i] = acc[ib][r];
    } else {
        if (i < out_dim) {
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[(16 * part + 4 * g + r) * out_dim + i] = acc[0][r];
        }
    }
}

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
    mlp_backward_kernel(uint32_t n, bool level_major, const float *__restrict__ enc, const float *__restrict__ dirs,
                        const float *__restrict__ params, const float *__restrict__ dL_ddrgbs,
                        float *__restrict__ dL_denc, float *__restrict__ workspace) {
    __shared__ float lds[L_END];
    __shared__ float scratch[kWaves * 2 * T_FLOATS];
    stage_weights(lds, params);
    const int lane = lane_id();
    const int i = lane & 15, g = lane >> 4;
    const int wv = threadIdx.x / kWave;
    float *zs = scratch + wv * 2 * T_FLOATS;
    float *ys = zs + T_FLOATS;
    const uint32_t n_tiles = (n + kTile - 1) / kTile;
    // the workgroup takes kWaves consecutive tiles at a time (a tile past the end
    // loads as zeros and contributes nothing), so all its waves meet at the barriers
    const uint32_t n_groups = (n_tiles + kWaves - 1) / kWaves;

    // this wave's quarter of each weight gradient (see dw_accumulate_part)
    v4f gW1[2], gW2[1], gW3[2], gW4[4], gW5[1];
#pragma unroll
    for (int k = 0; k < 2; ++k) gW1[k] = gW3[k] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    gW2[0] = gW5[0] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) gW4[k] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

    TileIn next = load_tile(enc, level_major, dirs, dL_ddrgbs, n, ((size_t)blockIdx.x * kWaves + wv) * kTile, i, g);
    for (uint32_t tg = blockIdx.x; tg < n_groups; tg += gridDim.x) {
        const size_t s0 = ((size_t)tg * kWaves + wv) * kTile;
        const size_t s = s0 + i;
        const bool live = s < n;
        const TileIn in = next;
        next = load_tile(enc, level_major, dirs, dL_ddrgbs, n, s0 + (size_t)gridDim.x * kWaves * kTile, i, g);
        Fwd f;
        f.x[0] = in.x[0];
        f.x[1] = in.x[1];
        forward_from_x(lds, in.dx, in.dy, in.dz, i, g, f);
        const v4f gin = in.gin;
        // colour head: d(pre-sigmoid), features 0..2 live in lanes g == 0
        v4f dr = {0.0f, 0.0f, 0.0f, 0.0f};
        float dsig = 0.0f;
        if (g == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float y = sigmoidf(f.rr[c]);
                dr[c] = gin[c + 1] * y * (1.0f - y);
            }
            dsig = gin[0] * det_expf(fminf(fmaxf(f.c[0][0], -15.0f), 15.0f));
        }
        // W5
        tile_to_lds<4>(zs, f.h3, i, g);
        tile_to_lds<1>(ys, &dr, i, g);
        __syncthreads();
        dw_accumulate_part<4, 1>(gW5, scratch, wv, i, g);
        v4f d3[4];
        layer_bwd<64, 16, S_W5>(lds + L_W5, &dr, d3, i, g);
        relu_mask<4>(d3, f.h3);
        __syncthreads();
        // W4
        tile_to_lds<4>(zs, f.h2, i, g);
        tile_to_lds<4>(ys, d3, i, g);
        __syncthreads();
        dw_accumulate_part<4, 4>(gW4, scratch, wv, i, g);
        v4f d2[4];
        layer_bwd<64, 64, S_W4>(lds + L_W4, d3, d2, i, g);
        relu_mask<4>(d2, f.h2);
        __syncthreads();
        // W3
        tile_to_lds<2>(zs, f.c, i, g);
        tile_to_lds<4>(ys, d2, i, g);
        __syncthreads();
        dw_accumulate_part<2, 4>(gW3, scratch, wv, i, g);
        v4f dc[2];
        layer_bwd<32, 64, S_W3>(lds + L_W3, d2, dc, i, g);
        if (g == 0) dc[0][0] += dsig;
        __syncthreads();
        // W2
        tile_to_lds<4>(zs, f.h1, i, g);
        tile_to_lds<1>(ys, &dc[0], i, g);
        __syncthreads();
        dw_accumulate_part<4, 1>(gW2, scratch, wv, i, g);
        v4f d1[4];
        layer_bwd<64, 16, S_W2>(lds + L_W2, &dc[0], d1, i, g);
        relu_mask<4>(d1, f.h1);
        __syncthreads();
        // W1
        tile_to_lds<2>(zs, f.x, i, g);
        tile_to_lds<4>(ys, d1, i, g);
        __syncthreads();
        dw_accumulate_part<2, 4>(gW1, scratch, wv, i, g);
        v4f dx[2];
        layer_bwd<32, 64, S_W1>(lds + L_W1, d1, dx, i, g);
        if (live) {
#pragma unroll
            for (int b = 0; b < 2; ++b) store_enc4(dL_denc, level_major, n, s, b, g, dx[b]);
        }
        __syncthreads();
    }

    // one workspace row per workgroup; each element written by one lane of one wave
    float *row = workspace + (size_t)blockIdx.x * P_END;
    dw_part_to_row<2, 4>(row + P_W1, gW1, 64, wv, i, g);
    dw_part_to_row<4, 1>(row + P_W2, gW2, 16, wv, i, g);
    dw_part_to_row<2, 4>(row + P_W3, gW3, 64, wv, i, g);
    dw_part_to_row<4, 4>(row + P_W4, gW4, 64, wv, i, g);
    dw_part_to_row<4, 1>(row + P_W5, gW5, 3, wv, i, g);
}

// out[p] = sum of workspace rows at column p.  Block: 64 columns x 16 waves; wave
// w sums rows w, w+16, ... with 8 loads in flight, then the waves' sums are added
// in wave order, so the result does not depend on scheduling.
constexpr int kSumWaves = 16;
__global__ void __launch_bounds__(kSumWaves * kWave)
    sum_rows_kernel(uint32_t rows, const float *__restrict__ workspace, float *__restrict__ out) {
    __shared__ float part[kSumWaves][kWave];
    const int lane = lane_id(), wave = (int)threadIdx.x / kWave;
    const uint32_t p = blockIdx.x * kWave + lane;
    const uint32_t pc = p < (uint32_t)P_END ? p : (uint32_t)P_END - 1;
    float acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0.0f;
    uint32_t r = wave;
    for (; r + 7 * kSumWaves < rows; r += 8 * kSumWaves) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += workspace[(size_t)(r + u * kSumWaves) * P_END + pc];
    }
    for (; r < rows; r += kSumWaves) acc[0] += workspace[(size_t)r * P_END + pc];
    part[wave][lane] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (wave == 0 && p < (uint32_t)P_END) {
        float s = 0.0f;
#pragma unroll
        for (int w = 0; w < kSumWaves; ++w) s += part[w][lane];
        out[p] = s;
    }
}

// ---------------------------------------------------------------------------
// f16 weights (the precision path): the same network with f16 weights and f16
// MFMA operands, f32 accumulation (v_mfma_f32_16x16x32_f16 for the layers,
// v_mfma_f32_16x16x16_f16 for the weight gradients), f32 activations between the
// products (ReLU and masks in f32, rounded to f16 only as MFMA inputs), f32
// density/sigmoid heads and f32 gradients out.  Cotangents arrive loss-scaled.
//
// K mapping.  A 16x16x32 product takes, on lane (i, g) = (lane % 16, lane / 16),
// eight K slots q = 0..7 of A's row i and B's column i.  Here slot q of lane group
// g stands for feature 32 kb + 4g + q (q < 4) or 32 kb + 16 + 4g + (q - 4)
// (q >= 4): exactly the registers lane (i, g) holds of two consecutive MFMA output
// tiles, so a layer's outputs are the next layer's B operand as they are, and the
// weights are read from LDS with the same mapping (two 8-byte runs per lane).
// ---------------------------------------------------------------------------

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// forward copies W^T [out][in + 8], backward copies W [in][out_pad + 8] (halves)
constexpr int FT1 = 40, FT2 = 72, FT3 = 40, FT4 = 72, FT5 = 72;
constexpr int H_F1 = 0;
constexpr int H_F2 = H_F1 + 64 * FT1;
constexpr int H_F3 = H_F2 + 16 * FT2;
constexpr int H_F4 = H_F3 + 64 * FT3;
constexpr int H_F5 = H_F4 + 64 * FT4;
constexpr int H_FEND = H_F5 + 16 * FT5;
constexpr int BT1 = 72, BT2 = 24, BT3 = 72, BT4 = 72, BT5 = 24;
constexpr int H_B1 = H_FEND;
constexpr int H_B2 = H_B1 + 32 * BT1;
constexpr int H_B3 = H_B2 + 64 * BT2;
constexpr int H_B4 = H_B3 + 32 * BT3;
constexpr int H_B5 = H_B4 + 64 * BT4;
constexpr int H_END = H_B5 + 64 * BT5;  // 24320 halves

__device__ __forceinline__ v4f mfma_h32(h8 a, h8 b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma_h16(h4 a, h4 b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ h8 pack_h8(v4f lo, v4f hi) {
    return h8{(_Float16)lo[0], (_Float16)lo[1], (_Float16)lo[2], (_Float16)lo[3],
              (_Float16)hi[0], (_Float16)hi[1], (_Float16)hi[2], (_Float16)hi[3]};
}

__device__ __forceinline__ h8 join_h4(h4 lo, h4 hi) {
    return h8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// Copies the flat f16 parameters into the forward (transposed) and, with
// kBackward, the backward LDS layouts; padding is zero.
template <bool kBackward>
__device__ void stage_weights_h(_Float16 *lds, const _Float16 *__restrict__ params) {
    const int end = kBackward ? H_END : H_FEND;
    for (int t = threadIdx.x; t < end; t += kThreads) lds[t] = (_Float16)0.0f;
    __syncthreads();
    // W1 [32][64], W2 [64][16], W3 [32][64], W4 [64][64], W5 [64][3] (flax [in][out])
    for (int t = threadIdx.x; t < 32 * 64; t += kThreads) {
        const int in = t / 64, out = t % 64;
        lds[H_F1 + out * FT1 + in] = params[P_W1 + t];
        lds[H_F3 + out * FT3 + in] = params[P_W3 + t];
        if (kBackward) {
            lds[H_B1 + in * BT1 + out] = params[P_W1 + t];
            lds[H_B3 + in * BT3 + out] = params[P_W3 + t];
        }
    }
    for (int t = threadIdx.x; t < 64 * 16; t += kThreads) {
        const int in = t / 16, out = t % 16;
        lds[H_F2 + out * FT2 + in] = params[P_W2 + t];
        if (kBackward) lds[H_B2 + in * BT2 + out] = params[P_W2 + t];
    }
    for (int t = threadIdx.x; t < 64 * 64; t += kThreads) {
        const int in = t / 64, out = t % 64;
        lds[H_F4 + out * FT4 + in] = params[P_W4 + t];
        if (kBackward) lds[H_B4 + in * BT4 + out] = params[P_W4 + t];
    }
    for (int t = threadIdx.x; t < 64 * 3; t += kThreads) {
        const int in = t / 3, out = t % 3;
        lds[H_F5 + out * FT5 + in] = params[P_W5 + t];
        if (kBackward) lds[H_B5 + in * BT5 + out] = params[P_W5 + t];
    }
    __syncthreads();
}

// Y^T = W^T Z^T with KB K-slices of 32 inputs and OB output tiles of 16.
template <int KB, int OB, int ST>
__device__ __forceinline__ void layer_fwd_h(const _Float16 *WT, const h8 *in, v4f *out, int i, int g) {
#pragma unroll
    for (int ob = 0; ob < OB; ++ob) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            const _Float16 *row = WT + (16 * ob + i) * ST + 32 * kb + 4 * g;
            const h8 a = join_h4(*reinterpret_cast<const h4 *>(row), *reinterpret_cast<const h4 *>(row + 16));
            acc = mfma_h32(a, in[kb], acc);
        }
        out[ob] = acc;
    }
}

// dZ^T = W dY^T with IB input tiles and KB K-slices of 32 outputs; kHalfK: the
// layer has 16 (or fewer) outputs, the upper half of the slice is zero.
template <int IB, int KB, int ST, bool kHalfK>
__device__ __forceinline__ void layer_bwd_h(const _Float16 *W, const h8 *dout, v4f *din, int i, int g) {
#pragma unroll
    for (int ib = 0; ib < IB; ++ib) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            const _Float16 *row = W + (16 * ib + i) * ST + 32 * kb + 4 * g;
            const h4 lo = *reinterpret_cast<const h4 *>(row);
            const h4 hi = kHalfK ? h4{0, 0, 0, 0} : *reinterpret_cast<const h4 *>(row + 16);
            acc = mfma_h32(join_h4(lo, hi), dout[kb], acc);
        }
        din[ib] = acc;
    }
}

struct FwdH {
    v4f x[2], h1[4], c[2], h2[4], h3[4], rr;  // f32 values; c[0] = o, c[1] = SH
};

__device__ __forceinline__ void forward_from_x_h(const _Float16 *lds, float dx, float dy, float dz, int i, int g,
                                                 FwdH &f) {
    const h8 xin = pack_h8(f.x[0], f.x[1]);
    layer_fwd_h<1, 4, FT1>(lds + H_F1, &xin, f.h1, i, g);
    relu<4>(f.h1);
    const h8 h1in[2] = {pack_h8(f.h1[0], f.h1[1]), pack_h8(f.h1[2], f.h1[3])};
    layer_fwd_h<2, 1, FT2>(lds + H_F2, h1in, &f.c[0], i, g);
    f.c[1] = sh4_part(dx, dy, dz, g);
    const h8 cin = pack_h8(f.c[0], f.c[1]);
    layer_fwd_h<1, 4, FT3>(lds + H_F3, &cin, f.h2, i, g);
    relu<4>(f.h2);
    const h8 h2in[2] = {pack_h8(f.h2[0], f.h2[1]), pack_h8(f.h2[2], f.h2[3])};
    layer_fwd_h<2, 4, FT4>(lds + H_F4, h2in, f.h3, i, g);
    relu<4>(f.h3);
    const h8 h3in[2] = {pack_h8(f.h3[0], f.h3[1]), pack_h8(f.h3[2], f.h3[3])};
    layer_fwd_h<2, 1, FT5>(lds + H_F5, h3in, &f.rr, i, g);
}


__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
    mlp_forward_h_kernel(uint32_t n, bool level_major, const float *__restrict__ enc, const float *__restrict__ dirs,
                         const _Float16 *__restrict__ params, float *__restrict__ drgbs) {
    __shared__ __align__(16) _Float16 lds[H_FEND];
    stage_weights_h<false>(lds, params);
    const int lane = lane_id();
    const int i = lane & 15, g = lane >> 4;
    const uint32_t n_tiles = (n + kTile - 1) / kTile;
    const uint32_t wave = blockIdx.x * kWaves + threadIdx.x / kWave;
    const uint32_t n_waves = gridDim.x * kWaves;
    TileIn next = load_tile(enc, level_major, dirs, nullptr, n, (size_t)min(wave, n_tiles - 1) * kTile, i, g);
    for (uint32_t t = wave; t < n_tiles; t += n_waves) {
        const size_t s0 = (size_t)t * kTile;
        const TileIn in = next;
        next = load_tile(enc, level_major, dirs, nullptr, n, (size_t)min(t + n_waves, n_tiles - 1) * kTile, i, g);
        FwdH f;
        f.x[0] = in.x[0];
        f.x[1] = in.x[1];
        forward_from_x_h(lds, in.dx, in.dy, in.dz, i, g, f);
        const size_t s = s0 + i;
        if (g == 0 && s < n) {
            const float sigma = det_expf(f.c[0][0]);
            v4f out = {sigma, sigmoidf(f.rr[0]), sigmoidf(f.rr[1]), sigmoidf(f.rr[2])};
            *reinterpret_cast<v4f *>(drgbs + 4 * s) = out;
        }
    }
}

// acc[ib*OB + ob] += Z^T dY over the tile's 16 samples (K = samples), f16 operands:
// lane (i, g) takes samples 4g..4g+3 of feature 16ib + i (A) and 16ob + i (B).
template <int IB, int OB>
__device__ __forceinline__ void dw_accumulate_h(v4f *acc, const float *zs, const float *ys, int i, int g) {
    h4 a[IB], bb[OB];
#pragma unroll
    for (int ib = 0; ib < IB; ++ib)
#pragma unroll
        for (int q = 0; q < 4; ++q) a[ib][q] = (_Float16)zs[(4 * g + q) * S_T + 16 * ib + i];
#pragma unroll
    for (int ob = 0; ob < OB; ++ob)
#pragma unroll
        for (int q = 0; q < 4; ++q) bb[ob][q] = (_Float16)ys[(4 * g + q) * S_T + 16 * ob + i];
#pragma unroll
    for (int ib = 0; ib < IB; ++ib)
#pragma unroll
        for (int ob = 0; ob < OB; ++ob) acc[ib * OB + ob] = mfma_h16(a[ib], bb[ob], acc[ib * OB + ob]);
}

// tile_to_lds for the f16 backward: the dW MFMAs take f16 operands, so the tile is
// rounded to f16 once here (the same rounding as at the MFMA) and the scratch
// takes half the LDS.
template <int NB>
__device__ __forceinline__ void tile_to_lds_h(_Float16 *scr, const v4f *z, int i, int g) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        h4 v;
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = (_Float16)z[b][r];
        *reinterpret_cast<h4 *>(scr + i * S_T + 16 * b + 4 * g) = v;
    }
}

// dw_accumulate_part with f16 operands (K = the tile's 16 samples per MFMA).
template <int IB, int OB>
__device__ __forceinline__ void dw_accumulate_part_h(v4f *acc, const _Float16 *scratch, int part, int i, int g) {
    static_assert(OB == 4 || (OB == 1 && IB == 4), "blocks split over four waves");
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const _Float16 *zs = scratch + w * 2 * T_FLOATS;
        const _Float16 *ys = zs + T_FLOATS;
        const int cb = OB == 4 ? 16 * part : 0;  // this wave's output block in ys
        h4 bb;
#pragma unroll
        for (int q = 0; q < 4; ++q) bb[q] = ys[(4 * g + q) * S_T + cb + i];
#pragma unroll
        for (int ib = 0; ib < (OB == 4 ? IB : 1); ++ib) {
            const int rb = OB == 4 ? 16 * ib : 16 * part;  // input block in zs
            h4 a;
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = zs[(4 * g + q) * S_T + rb + i];
            acc[ib] = mfma_h16(a, bb, acc[ib]);
        }
    }
}

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
    mlp_backward_h_kernel(uint32_t n, bool level_major, const float *__restrict__ enc, const float *__restrict__ dirs,
                          const _Float16 *__restrict__ params, const float *__restrict__ dL_ddrgbs,
                          float *__restrict__ dL_denc, float *__restrict__ workspace) {
    __shared__ __align__(16) _Float16 lds[H_END];
    __shared__ __align__(16) _Float16 scratch[kWaves * 2 * T_FLOATS];
    stage_weights_h<true>(lds, params);
    const int lane = lane_id();
    const int i = lane & 15, g = lane >> 4;
    const int wv = threadIdx.x / kWave;
    _Float16 *zs = scratch + wv * 2 * T_FLOATS;
    _Float16 *ys = zs + T_FLOATS;
    const uint32_t n_tiles = (n + kTile - 1) / kTile;
    // kWaves consecutive tiles per workgroup step; tiles past the end load as zeros
    const uint32_t n_groups = (n_tiles + kWaves - 1) / kWaves;
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};

    // this wave's quarter of each weight gradient (see dw_accumulate_part)
    v4f gW1[2], gW2[1], gW3[2], gW4[4], gW5[1];
#pragma unroll
    for (int k = 0; k < 2; ++k) gW1[k] = gW3[k] = zero;
    gW2[0] = gW5[0] = zero;
#pragma unroll
    for (int k = 0; k < 4; ++k) gW4[k] = zero;

    TileIn next = load_tile(enc, level_major, dirs, dL_ddrgbs, n, ((size_t)blockIdx.x * kWaves + wv) * kTile, i, g);
    for (uint32_t tg = blockIdx.x; tg < n_groups; tg += gridDim.x) {
        const size_t s0 = ((size_t)tg * kWaves + wv) * kTile;
        const size_t s = s0 + i;
        const bool live = s < n;
        const TileIn in = next;
        next = load_tile(enc, level_major, dirs, dL_ddrgbs, n, s0 + (size_t)gridDim.x * kWaves * kTile, i, g);
        FwdH f;
        f.x[0] = in.x[0];
        f.x[1] = in.x[1];
        forward_from_x_h(lds, in.dx, in.dy, in.dz, i, g, f);
        const v4f gin = in.gin;
        v4f dr = zero;
        float dsig = 0.0f;
        if (g == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float y = sigmoidf(f.rr[c]);
                dr[c] = gin[c + 1] * y * (1.0f - y);
            }
            dsig = gin[0] * det_expf(fminf(fmaxf(f.c[0][0], -15.0f), 15.0f));
        }
        // W5
        tile_to_lds_h<4>(zs, f.h3, i, g);
        tile_to_lds_h<1>(ys, &dr, i, g);
        __syncthreads();
        dw_accumulate_part_h<4, 1>(gW5, scratch, wv, i, g);
        const h8 drin = pack_h8(dr, zero);
        v4f d3[4];
        layer_bwd_h<4, 1, BT5, true>(lds + H_B5, &drin, d3, i, g);
        relu_mask<4>(d3, f.h3);
        __syncthreads();
        // W4
        tile_to_lds_h<4>(zs, f.h2, i, g);
        tile_to_lds_h<4>(ys, d3, i, g);
        __syncthreads();
        dw_accumulate_part_h<4, 4>(gW4, scratch, wv, i, g);
        const h8 d3in[2] = {pack_h8(d3[0], d3[1]), pack_h8(d3[2], d3[3])};
        v4f d2[4];
        layer_bwd_h<4, 2, BT4, false>(lds + H_B4, d3in, d2, i, g);
        relu_mask<4>(d2, f.h2);
        __syncthreads();
        // W3
        tile_to_lds_h<2>(zs, f.c, i, g);
        tile_to_lds_h<4>(ys, d2, i, g);
        __syncthreads();
        dw_accumulate_part_h<2, 4>(gW3, scratch, wv, i, g);
        const h8 d2in[2] = {pack_h8(d2[0], d2[1]), pack_h8(d2[2], d2[3])};
        v4f dc[2];
        layer_bwd_h<2, 2, BT3, false>(lds + H_B3, d2in, dc, i, g);
        if (g == 0) dc[0][0] += dsig;
        __syncthreads();
        // W2
        tile_to_lds_h<4>(zs, f.h1, i, g);
        tile_to_lds_h<1>(ys, &dc[0], i, g);
        __syncthreads();
        dw_accumulate_part_h<4, 1>(gW2, scratch, wv, i, g);
        const h8 dcin = pack_h8(dc[0], zero);
        v4f d1[4];
        layer_bwd_h<4, 1, BT2, true>(lds + H_B2, &dcin, d1, i, g);
        relu_mask<4>(d1, f.h1);
        __syncthreads();
        // W1
        tile_to_lds_h<2>(zs, f.x, i, g);
        tile_to_lds_h<4>(ys, d1, i, g);
        __syncthreads();
        dw_accumulate_part_h<2, 4>(gW1, scratch, wv, i, g);
        const h8 d1in[2] = {pack_h8(d1[0], d1[1]), pack_h8(d1[2], d1[3])};
        v4f dx[2];
        layer_bwd_h<2, 2, BT1, false>(lds + H_B1, d1in, dx, i, g);
        if (live) {
#pragma unroll
            for (int b = 0; b < 2; ++b) store_enc4(dL_denc, level_major, n, s, b, g, dx[b]);
        }
        __syncthreads();
    }

    // one workspace row per workgroup; each element written by one lane of one wave
    float *row = workspace + (size_t)blockIdx.x * P_END;
    dw_part_to_row<2, 4>(row + P_W1, gW1, 64, wv, i, g);
    dw_part_to_row<4, 1>(row + P_W2, gW2, 16, wv, i, g);
    dw_part_to_row<2, 4>(row + P_W3, gW3, 64, wv, i, g);
    dw_part_to_row<4, 4>(row + P_W4, gW4, 64, wv, i, g);
    dw_part_to_row<4, 1>(row + P_W5, gW5, 3, wv, i, g);
}



// ---------------------------------------------------------------------------
// Fused field query of the inference renderer: hash encoding (L = 16, F = 2) and
// MLP in one pass over the [slots, cap] samples of one ray-pool round, of which
// slot j holds n_samples[j]; padding entries get drgbs = 0.  Lane (i, g) of a tile
// computes exactly the four levels 2g, 2g+1, 8+2g, 9+2g of sample i that its MFMA
// B operand needs (features 4g..4g+3 and 16+4g..16+4g+3), with the encoder's
// arithmetic, so the result equals hashgrid_encode followed by nerf_mlp_forward.
// ---------------------------------------------------------------------------

template <typename TT>
__device__ __forceinline__ float2 encode_level(const HashGridDescriptor &d, const LevelTables &lt,
                                               const TT *__restrict__ table, float px, float py, float pz,
                                               uint32_t l) {
    uint32_t idx[8];
    float w[8];
    level_vertices(px, py, pz, lt.scale[l], lt.res[l], l < d.first_hash_level, d.T, (d.T & (d.T - 1u)) == 0,
                   lt.offset[l], d.n_rows - 1u, idx, w);
    float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const TT *row = table + (size_t)idx[c] * 2;
        a0 = a0 + to_f32<TT>(row[0]) * w[c];
        a1 = a1 + to_f32<TT>(row[1]) * w[c];
    }
    return make_float2(a0, a1);
}

template <typename TT, bool kHalf>
__global__ void __launch_bounds__(kThreads)
    field_chunks_kernel(HashGridDescriptor d, uint32_t n_slots, uint32_t cap, uint32_t n_total,
                        const float *__restrict__ xyzs, const uint32_t *__restrict__ n_samples,
                        const uint32_t *__restrict__ indices, const float *__restrict__ rays_d,
                        const TT *__restrict__ table, const void *__restrict__ params, float *__restrict__ drgbs) {
    __shared__ __align__(16) float lds_f[kHalf ? 1 : L_END];
    __shared__ __align__(16) _Float16 lds_h[kHalf ? H_FEND : 1];
    __shared__ LevelTables lt;
    if (kHalf) stage_weights_h<false>(lds_h, static_cast<const _Float16 *>(params));
    else stage_weights(lds_f, static_cast<const float *>(params));
    if (threadIdx.x < 16) {
        lt.scale[threadIdx.x] = d.scales[threadIdx.x];
        lt.res[threadIdx.x] = d.resolutions[threadIdx.x];
        lt.offset[threadIdx.x] = d.offsets[threadIdx.x];
    }
    __syncthreads();
    const int lane = lane_id();
    const int i = lane & 15, g = lane >> 4;
    const uint64_t n_entries = (uint64_t)n_slots * cap;
    const uint64_t n_tiles = (n_entries + kTile - 1) / kTile;
    const uint32_t wave = blockIdx.x * kWaves + threadIdx.x / kWave;
    const uint32_t n_waves = gridDim.x * kWaves;
    const Divisor two_b = make_divisor(2.0f * d.bound);
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint64_t t = wave; t < n_tiles; t += n_waves) {
        const uint64_t e = t * kTile + i;
        bool live = false;
        uint32_t ray = 0;
        if (e < n_entries) {
            const uint32_t slot = (uint32_t)(e / cap), k = (uint32_t)(e % cap);
            live = k < n_samples[slot];
            if (live) {
                ray = indices[slot];
                live = ray < n_total;
            }
        }
        if (__ballot(live) == 0ull) {
            if (g == 0 && e < n_entries) *reinterpret_cast<v4f *>(drgbs + 4 * e) = zero;
            continue;
        }
        float px = 0.0f, py = 0.0f, pz = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
        if (live) {
            px = divide(two_b, xyzs[3 * e + 0] + d.bound);
            py = divide(two_b, xyzs[3 * e + 1] + d.bound);
            pz = divide(two_b, xyzs[3 * e + 2] + d.bound);
            dx = rays_d[3 * (size_t)ray + 0];
            dy = rays_d[3 * (size_t)ray + 1];
            dz = rays_d[3 * (size_t)ray + 2];
        }
        v4f x[2];
        if (live) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const float2 lo = encode_level(d, lt, table, px, py, pz, 8 * b + 2 * g);
                const float2 hi = encode_level(d, lt, table, px, py, pz, 8 * b + 2 * g + 1);
                x[b] = v4f{lo.x, lo.y, hi.x, hi.y};
            }
        } else {
            x[0] = x[1] = zero;
        }
        v4f out = zero;
        float o0;
        v4f rr;
        if (kHalf) {
            FwdH f;
            f.x[0] = x[0];
            f.x[1] = x[1];
            forward_from_x_h(lds_h, dx, dy, dz, i, g, f);
            o0 = f.c[0][0];
            rr = f.rr;
        } else {
            Fwd f;
            f.x[0] = x[0];
            f.x[1] = x[1];
            forward_from_x(lds_f, dx, dy, dz, i, g, f);
            o0 = f.c[0][0];
            rr = f.rr;
        }
        if (live) out = v4f{det_expf(o0), sigmoidf(rr[0]), sigmoidf(rr[1]), sigmoidf(rr[2])};
        if (g == 0 && e < n_entries) *reinterpret_cast<v4f *>(drgbs + 4 * e) = out;
    }
}

// ---------------------------------------------------------------------------
// Occupancy-grid update (utils/types.py:1149-1225).  A wavefront takes 16 chosen
// cells; lane (i, g) derives cell i's random point, computes the four levels of
// its encoding that its MFMA operand needs (as in field_chunks_kernel), and the
// density MLP (32 -> 64 -> 16, nerfs.py:66-77) runs on MFMA; lane (i, 0) then
// applies trunc_exp and the max.  Cell choice and points come from draws keyed by
// (seed, step, cascade << 32 | slot).
// ---------------------------------------------------------------------------

constexpr uint32_t kDecayBits = 0x3f733333u;  // 0.95f (types.py:1163)

struct OgridArgs {
    uint32_t G, cascade, n_alive, n_update, half;  // half: slots of the uniform first half
    bool update_all;
    uint64_t seed;
    uint32_t step;
    float bound, mip, hcw;
};

// jnp.maximum: NaN if either operand is NaN
__device__ __forceinline__ float max_nan(float a, float b) { return (a != a || b != b) ? a + b : fmaxf(a, b); }

// Position of bit k (0-based, k < popcount(w)) among the set bits of w.
__device__ __forceinline__ uint32_t select_bit(uint64_t w, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (int width = 32; width > 0; width >>= 1) {
        const uint64_t low = w & ((1ull << width) - 1ull);
        const uint32_t c = (uint32_t)__popcll(low);
        if (k >= c) {
            k -= c;
            pos += width;
            w >>= width;
        } else {
            w = low;
        }
    }
    return pos;
}

// Morton index (within the cascade) of update slot j.
__device__ __forceinline__ uint32_t ogrid_cell(const OgridArgs &a, uint32_t j, const uint32_t *__restrict__ alive,
                                               const uint32_t *__restrict__ prefix, const uint64_t *__restrict__ words) {
    if (a.update_all) return alive[j];
    const uint64_t key = ((uint64_t)a.cascade << 32) | j;
    if (j >= a.half) {
        const uint32_t W = a.G * a.G * a.G / 64u;  // 64-cell words per cascade (G >= 4)
        const uint32_t n_occ = prefix[W];
        if (n_occ > 0) {
            const uint32_t r = below(draw(a.seed, a.step, key, 1), n_occ);
            uint32_t lo = 0, hi = W;  // largest w with prefix[w] <= r
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (prefix[mid] <= r) lo = mid;
                else hi = mid;
            }
            return lo * 64u + select_bit(words[(size_t)a.cascade * W + lo], r - prefix[lo]);
        }
    }
    return alive[below(draw(a.seed, a.step, key, 0), a.n_alive)];
}

// Random point of cell m (types.py:1193-1207): the cell centre mapped into
// [-mip + hcw, mip - hcw], plus a uniform offset in [-hcw, hcw) per axis from three
// 21-bit uniforms of one draw.
__device__ __forceinline__ void ogrid_point(const OgridArgs &a, uint32_t m, uint32_t j, float &x, float &y,
                                            float &z) {
    const uint64_t r = draw(a.seed, a.step, ((uint64_t)a.cascade << 32) | j, 2);
    const float u[3] = {(float)((uint32_t)(r >> 43) & 0x1fffffu) * (1.0f / 2097152.0f),
                        (float)((uint32_t)(r >> 22) & 0x1fffffu) * (1.0f / 2097152.0f),
                        (float)((uint32_t)(r >> 1) & 0x1fffffu) * (1.0f / 2097152.0f)};
    const uint32_t v[3] = {morton_compact(m), morton_compact(m >> 1), morton_compact(m >> 2)};
    const float g1 = (float)(a.G - 1u), span = a.mip - a.hcw;
    float p[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = ((float)v[k] / g1) * 2.0f - 1.0f;
        p[k] = c * span + (u[k] * 2.0f - 1.0f) * a.hcw;
    }
    x = p[0];
    y = p[1];
    z = p[2];
}

template <typename TT, bool kHalf>
__global__ void __launch_bounds__(kThreads)
    ogrid_update_kernel(HashGridDescriptor d, OgridArgs a, const uint32_t *__restrict__ alive,
                        const uint32_t *__restrict__ prefix, const uint64_t *__restrict__ words,
                        const TT *__restrict__ table, const void *__restrict__ params, float *__restrict__ density) {
    __shared__ __align__(16) float lds_f[kHalf ? 1 : L_END];
    __shared__ __align__(16) _Float16 lds_h[kHalf ? H_FEND : 1];
    __shared__ LevelTables lt;
    if (kHalf) stage_weights_h<false>(lds_h, static_cast<const _Float16 *>(params));
    else stage_weights(lds_f, static_cast<const float *>(params));
    if (threadIdx.x < 16) {
        lt.scale[threadIdx.x] = d.scales[threadIdx.x];
        lt.res[threadIdx.x] = d.resolutions[threadIdx.x];
        lt.offset[threadIdx.x] = d.offsets[threadIdx.x];
    }
    __syncthreads();
    const int lane = lane_id();
    const int i = lane & 15, g = lane >> 4;
    const uint32_t n_tiles = (a.n_update + kTile - 1) / kTile;
    const uint32_t wave = blockIdx.x * kWaves + threadIdx.x / kWave;
    const uint32_t n_waves = gridDim.x * kWaves;
    const Divisor two_b = make_divisor(2.0f * a.bound);
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t G3 = a.G * a.G * a.G;
    float *dens = density + (size_t)a.cascade * G3;
    for (uint32_t t = wave; t < n_tiles; t += n_waves) {
        const uint32_t j = t * kTile + i;
        bool live = j < a.n_update;
        uint32_t m = 0;
        v4f x[2] = {zero, zero};
        if (live) {
            m = ogrid_cell(a, j, alive, prefix, words);
            live = m < G3;  // an alive index past the cascade is dropped, as JAX drops it
        }
        if (live) {
            float wx, wy, wz;
            ogrid_point(a, m, j, wx, wy, wz);
            const float px = divide(two_b, wx + a.bound), py = divide(two_b, wy + a.bound),
                        pz = divide(two_b, wz + a.bound);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const float2 lo = encode_level(d, lt, table, px, py, pz, 8 * b + 2 * g);
                const float2 hi = encode_level(d, lt, table, px, py, pz, 8 * b + 2 * g + 1);
                x[b] = v4f{lo.x, lo.y, hi.x, hi.y};
            }
        }
        float o0;
        if (kHalf) {
            FwdH f;
            const h8 xin = pack_h8(x[0], x[1]);
            layer_fwd_h<1, 4, FT1>(lds_h + H_F1, &xin, f.h1, i, g);
            relu<4>(f.h1);
            const h8 h1in[2] = {pack_h8(f.h1[0], f.h1[1]), pack_h8(f.h1[2], f.h1[3])};
            layer_fwd_h<2, 1, FT2>(lds_h + H_F2, h1in, &f.c[0], i, g);
            o0 = f.c[0][0];
        } else {
            Fwd f;
            layer_fwd<32, 64, S_W1>(lds_f + L_W1, x, f.h1, i, g);
            relu<4>(f.h1);
            layer_fwd<64, 16, S_W2>(lds_f + L_W2, f.h1, &f.c[0], i, g);
            o0 = f.c[0][0];
        }
        if (g == 0 && live) {
            const float sigma = det_expf(o0);
            if (a.update_all) {
                // each alive cell is in exactly one slot: decay and max in place
                dens[m] = max_nan(dens[m] * __uint_as_float(kDecayBits), sigma);
            } else {
                // decayed by ogrid_decay_kernel; densities of alive cells are >= +0 (or
                // NaN), so the unsigned order of their bits is the float order, NaN on top
                atomicMax(reinterpret_cast<unsigned int *>(dens + m), __float_as_uint(sigma));
            }
        }
    }
}

__global__ void __launch_bounds__(256) ogrid_decay_kernel(uint32_t n_alive, uint32_t G3, const uint32_t *__restrict__ alive,
                                                          float *__restrict__ dens) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= n_alive) return;
    const uint32_t m = alive[j];
    if (m < G3) dens[m] = dens[m] * __uint_as_float(kDecayBits);
}

// prefix[w] = occupied cells in the cascade's 64-cell words [0, w), w = 0..W
// (one workgroup; each thread counts a run of consecutive words).
constexpr int kPrefixThreads = 1024;
__global__ void __launch_bounds__(kPrefixThreads)
    ogrid_prefix_kernel(uint32_t W, const uint64_t *__restrict__ words, uint32_t *__restrict__ prefix) {
    const uint32_t per = (W + kPrefixThreads - 1) / kPrefixThreads;
    const uint32_t w0 = min(W, threadIdx.x * per), w1 = min(W, w0 + per);
    uint32_t own = 0;
    for (uint32_t w = w0; w < w1; ++w) own += (uint32_t)__popcll(words[w]);
    uint32_t total = 0;
    uint32_t run = block_exclusive_scan<kPrefixThreads>(own, &total);
    for (uint32_t w = w0; w < w1; ++w) {
        prefix[w] = run;
        run += (uint32_t)__popcll(words[w]);
    }
    if (threadIdx.x == 0) prefix[W] = total;
}

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

uint32_t forward_grid(uint32_t n) {
    const uint32_t tiles = (n + kTile - 1) / kTile;
    return std::max<uint32_t>(1, std::min<uint32_t>((tiles + kWaves - 1) / kWaves, 1024));
}

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void nerf_mlp_forward(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "nerf_mlp_forward";
    NerfMlpDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.enc_layout > NGP_ENC_LEVEL_MAJOR || d.param_dtype > NGP_DTYPE_F16) {
        set_error(kBadDescriptor, "%s: enc_layout=%u and param_dtype=%u must be 0 or 1", target, d.enc_layout,
                  d.param_dtype);
        return;
    }
    if (d.n_samples == 0) return;
    if (!check_buffers(buffers, 4, target)) return;
    if (!aligned16(buffers[0]) || !aligned16(buffers[3])) {
        set_error(kBadArgument, "%s: encodings and drgbs must be 16-byte aligned", target);
        return;
    }
    const bool lm = d.enc_layout == NGP_ENC_LEVEL_MAJOR;
    if (d.param_dtype == NGP_DTYPE_F16) {
        hipLaunchKernelGGL(mlp_forward_h_kernel, dim3(forward_grid(d.n_samples)), dim3(kThreads), 0,
                           (hipStream_t)stream_, d.n_samples, lm, (const float *)buffers[0], (const float *)buffers[1],
                           (const _Float16 *)buffers[2], (float *)buffers[3]);
    } else {
        hipLaunchKernelGGL(mlp_forward_kernel, dim3(forward_grid(d.n_samples)), dim3(kThreads), 0, (hipStream_t)stream_,
                           d.n_samples, lm, (const float *)buffers[0], (const float *)buffers[1],
                           (const float *)buffers[2], (float *)buffers[3]);
    }
    check_launch(target);
}

namespace {

// Checks the descriptor and the six data buffers (encodings, dirs, params,
// dL_ddrgbs, dL_dencodings, dL_dparams) shared by both backward entry points.
bool check_mlp_backward(const NerfMlpDescriptor &d, void **buffers, const char *target) {
    if (d.n_workspace_blocks == 0 || d.enc_layout > NGP_ENC_LEVEL_MAJOR || d.param_dtype > NGP_DTYPE_F16) {
        set_error(kBadDescriptor, "%s: n_workspace_blocks must be positive, enc_layout and param_dtype 0 or 1",
                  target);
        return false;
    }
    if (!check_buffer_list(buffers, {2, 5}, target)) return false;
    if (d.n_samples > 0 && !check_buffers(buffers, 6, target)) return false;
    if (d.n_samples > 0 && (!aligned16(buffers[0]) || !aligned16(buffers[3]) || !aligned16(buffers[4]))) {
        set_error(kBadArgument, "%s: encodings, dL_ddrgbs and dL_dencodings must be 16-byte aligned", target);
        return false;
    }
    return true;
}

// The backward kernel (one workspace row per workgroup, every element of the row
// written) and the fixed-order row sum into dL_dparams.
void launch_mlp_backward(const NerfMlpDescriptor &d, hipStream_t stream, void **buffers, float *workspace,
                         const char *target) {
    const bool lm = d.enc_layout == NGP_ENC_LEVEL_MAJOR;
    if (d.param_dtype == NGP_DTYPE_F16) {
        hipLaunchKernelGGL(mlp_backward_h_kernel, dim3(d.n_workspace_blocks), dim3(kThreads), 0, stream, d.n_samples,
                           lm, (const float *)buffers[0], (const float *)buffers[1], (const _Float16 *)buffers[2],
                           (const float *)buffers[3], (float *)buffers[4], workspace);
    } else {
        hipLaunchKernelGGL(mlp_backward_kernel, dim3(d.n_workspace_blocks), dim3(kThreads), 0, stream, d.n_samples, lm,
                           (const float *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                           (const float *)buffers[3], (float *)buffers[4], workspace);
    }
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(sum_rows_kernel, dim3(div_up(P_END, kWave)), dim3(kSumWaves * kWave), 0, stream,
                       d.n_workspace_blocks, (const float *)workspace, (float *)buffers[5]);
    check_launch(target);
}

// Workspace of nerf_mlp_backward_no_alloc: the weight-gradient rows of up to
// NGP_MLP_NO_ALLOC_MAX_BLOCKS workgroups (512 x 9408 f32 = 18.4 MiB per GPU, held
// until the process ends).  Host code must not pass `scratch` itself to a kernel:
// on the host the name is a stand-in, not the device address.  The address of each
// device's copy comes from hipGetSymbolAddress, looked up once per device.
static __device__ __align__(256) float scratch[(size_t)NGP_MLP_NO_ALLOC_MAX_BLOCKS * P_END];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float *g_scratch_ptr[kMaxDevices] = {nullptr};

// The address of `scratch` on the device that owns `stream` (the current device for
// the null stream), looked up with that device current on its first use.
float *scratch_for(hipStream_t stream, const char *target) {
    int dev = -1;
    hipError_t e = stream == nullptr ? hipGetDevice(&dev) : hipStreamGetDevice(stream, &dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error(kBadArgument, "%s: cannot determine the stream's device (%s, device %d)", target,
                  hipGetErrorString(e), dev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] != nullptr) return g_scratch_ptr[dev];
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) {
        set_error(kLaunchFailed, "%s: hipGetDevice failed", target);
        return nullptr;
    }
    if (prev != dev && hipSetDevice(dev) != hipSuccess) {
        set_error(kLaunchFailed, "%s: hipSetDevice(%d) failed", target, dev);
        return nullptr;
    }
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != dev) (void)hipSetDevice(prev);
    if (e != hipSuccess || ptr == nullptr) {
        set_error(kLaunchFailed, "%s: hipGetSymbolAddress(scratch) failed on device %d: %s", target, dev,
                  hipGetErrorString(e));
        return nullptr;
    }
    g_scratch_ptr[dev] = (float *)ptr;
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" void nerf_mlp_backward(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "nerf_mlp_backward";
    NerfMlpDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!check_mlp_backward(d, buffers, target)) return;
    if (!check_buffer_list(buffers, {6}, target)) return;
    launch_mlp_backward(d, (hipStream_t)stream_, buffers, (float *)buffers[6], target);
}

extern "C" void nerf_mlp_backward_no_alloc(ngp_stream_t stream_, void **buffers, const char *opaque,
                                           size_t opaque_len) {
    const char *target = "nerf_mlp_backward_no_alloc";
    NerfMlpDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (!check_mlp_backward(d, buffers, target)) return;
    if (d.n_workspace_blocks > NGP_MLP_NO_ALLOC_MAX_BLOCKS) {
        set_error(kBadDescriptor, "%s: n_workspace_blocks=%u exceeds the fixed scratch of %u workgroup rows", target,
                  d.n_workspace_blocks, (unsigned)NGP_MLP_NO_ALLOC_MAX_BLOCKS);
        return;
    }
    hipStream_t stream = (hipStream_t)stream_;
    float *ws = scratch_for(stream, target);
    if (ws == nullptr) return;
    launch_mlp_backward(d, stream, buffers, ws, target);
}

namespace {
// The NeRF position encoder's descriptor: L = 16, F = 2, a consistent level table.
bool check_nerf_grid(const HashGridDescriptor &g, uint32_t param_dtype, const char *target) {
    if (g.L != 16 || g.F != 2 || g.T == 0 || g.n_rows == 0 || g.first_hash_level > g.L || g.offsets[0] != 0 ||
        g.offsets[g.L] != g.n_rows || !(g.bound > 0.0f) || g.table_dtype > NGP_DTYPE_F16 ||
        param_dtype > NGP_DTYPE_F16) {
        set_error(kBadDescriptor, "%s: needs the NeRF grid (L=16, F=2) with a consistent level table and "
                  "f32/f16 dtypes", target);
        return false;
    }
    for (uint32_t l = 0; l < g.L; ++l) {
        if (g.offsets[l + 1] < g.offsets[l]) {
            set_error(kBadDescriptor, "%s: offsets must be non-decreasing", target);
            return false;
        }
    }
    return true;
}
}  // namespace

extern "C" void nerf_field_chunks(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "nerf_field_chunks";
    NerfFieldChunksDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    const HashGridDescriptor &g = d.grid;
    if (!check_nerf_grid(g, d.param_dtype, target)) return;
    if (d.cap == 0) {
        set_error(kBadDescriptor, "%s: cap must be positive", target);
        return;
    }
    if (d.n_slots == 0) return;
    if (!check_buffers(buffers, 7, target)) return;
    if (!aligned16(buffers[6])) {
        set_error(kBadArgument, "%s: drgbs must be 16-byte aligned", target);
        return;
    }
    const uint64_t tiles = ((uint64_t)d.n_slots * d.cap + kTile - 1) / kTile;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((tiles + kWaves - 1) / kWaves, 2048));
    const bool t16 = g.table_dtype == NGP_DTYPE_F16, w16 = d.param_dtype == NGP_DTYPE_F16;
    hipStream_t s = (hipStream_t)stream;
    const float *xyzs = (const float *)buffers[0];
    const uint32_t *ns = (const uint32_t *)buffers[1];
    const uint32_t *idx = (const uint32_t *)buffers[2];
    const float *rd = (const float *)buffers[3];
    const void *params = buffers[5];
    float *out = (float *)buffers[6];
    if (t16 && w16)
        hipLaunchKernelGGL((field_chunks_kernel<_Float16, true>), dim3(grid), dim3(kThreads), 0, s, g, d.n_slots, d.cap,
                           d.n_total_rays, xyzs, ns, idx, rd, (const _Float16 *)buffers[4], params, out);
    else if (t16)
        hipLaunchKernelGGL((field_chunks_kernel<_Float16, false>), dim3(grid), dim3(kThreads), 0, s, g, d.n_slots,
                           d.cap, d.n_total_rays, xyzs, ns, idx, rd, (const _Float16 *)buffers[4], params, out);
    else if (w16)
        hipLaunchKernelGGL((field_chunks_kernel<float, true>), dim3(grid), dim3(kThreads), 0, s, g, d.n_slots, d.cap,
                           d.n_total_rays, xyzs, ns, idx, rd, (const float *)buffers[4], params, out);
    else
        hipLaunchKernelGGL((field_chunks_kernel<float, false>), dim3(grid), dim3(kThreads), 0, s, g, d.n_slots, d.cap,
                           d.n_total_rays, xyzs, ns, idx, rd, (const float *)buffers[4], params, out);
    check_launch(target);
}

extern "C" void occupancy_grid_update(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "occupancy_grid_update";
    OccupancyGridUpdateDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    const HashGridDescriptor &g = d.grid;
    if (!check_nerf_grid(g, d.param_dtype, target)) return;
    if (d.G < 4 || d.G > 1024 || (d.G & (d.G - 1u)) != 0 || d.K == 0 || d.cascade >= d.K || d.update_all > 1 ||
        (uint64_t)d.n_alive > (uint64_t)d.G * d.G * d.G) {
        set_error(kBadDescriptor, "%s: G must be a power of two in [4, 1024], cascade < K, update_all 0 or 1, "
                  "n_alive <= G^3", target);
        return;
    }
    if (d.n_alive == 0) return;  // no alive cell: nothing to decay or update
    if (!check_buffers(buffers, 5, target)) return;
    if (!d.update_all && !check_buffer_list(buffers, {5}, target)) return;
    if ((reinterpret_cast<uintptr_t>(buffers[1]) & 7u) != 0) {
        set_error(kBadArgument, "%s: the occupancy bitfield must be 8-byte aligned", target);
        return;
    }
    hipStream_t s = (hipStream_t)stream;
    OgridArgs a;
    a.G = d.G;
    a.cascade = d.cascade;
    a.n_alive = d.n_alive;
    a.update_all = d.update_all != 0;
    a.seed = ((uint64_t)d.seed_hi << 32) | d.seed_lo;
    a.step = d.step;
    a.bound = g.bound;
    a.mip = std::fmin(g.bound, std::ldexp(1.0f, (int)std::min<uint32_t>(d.cascade, 126u)));
    a.hcw = a.mip / (float)d.G;
    const uint32_t *alive = (const uint32_t *)buffers[0];
    const uint64_t *words = (const uint64_t *)buffers[1];
    float *density = (float *)buffers[4];
    uint32_t *prefix = (uint32_t *)buffers[5];
    const uint32_t W = d.G * d.G * d.G / 64u;  // 64-cell words per cascade
    if (a.update_all) {
        a.n_update = d.n_alive;
        a.half = 0;
        prefix = nullptr;
    } else {
        const uint32_t M = std::max<uint32_t>(1, d.n_alive / 2), h = std::max<uint32_t>(1, M / 2);
        a.n_update = 2 * h;
        a.half = h;
        float *dens = density + (size_t)d.cascade * d.G * d.G * d.G;
        hipLaunchKernelGGL(ogrid_decay_kernel, dim3(div_up(d.n_alive, 256)), dim3(256), 0, s, d.n_alive,
                           d.G * d.G * d.G, alive, dens);
        if (!check_launch(target)) return;
        hipLaunchKernelGGL(ogrid_prefix_kernel, dim3(1), dim3(kPrefixThreads), 0, s, W, words + (size_t)d.cascade * W,
                           prefix);
        if (!check_launch(target)) return;
    }
    const uint32_t tiles = div_up(a.n_update, kTile);
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(div_up(tiles, kWaves), 2048));
    const void *params = buffers[3];
    const bool t16 = g.table_dtype == NGP_DTYPE_F16, w16 = d.param_dtype == NGP_DTYPE_F16;
    if (t16 && w16)
        hipLaunchKernelGGL((ogrid_update_kernel<_Float16, true>), dim3(grid), dim3(kThreads), 0, s, g, a, alive,
                           prefix, words, (const _Float16 *)buffers[2], params, density);
    else if (t16)
        hipLaunchKernelGGL((ogrid_update_kernel<_Float16, false>), dim3(grid), dim3(kThreads), 0, s, g, a, alive,
                           prefix, words, (const _Float16 *)buffers[2], params, density);
    else if (w16)
        hipLaunchKernelGGL((ogrid_update_kernel<float, true>), dim3(grid), dim3(kThreads), 0, s, g, a, alive, prefix,
                           words, (const float *)buffers[2], params, density);
    else
        hipLaunchKernelGGL((ogrid_update_kernel<float, false>), dim3(grid), dim3(kThreads), 0, s, g, a, alive, prefix,
                           words, (const float *)buffers[2], params, density);
    check_launch(target);
}
