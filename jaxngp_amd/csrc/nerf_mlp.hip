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
#include "ngp_common.h"

namespace ngp {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kThreads = kWaves * kWave;
constexpr int kTile = 16;

// flax parameter offsets (floats) in the flat parameter array
constexpr int P_W1 = 0, P_W2 = 2048, P_W3 = 3072, P_W4 = 5120, P_W5 = 9216, P_END = 9408;
static_assert(P_END == NGP_MLP_N_PARAMS, "parameter count");

// LDS copies: W[in][out_padded + 1] (odd row stride)
constexpr int S_W1 = 65, S_W2 = 17, S_W3 = 65, S_W4 = 65, S_W5 = 17;
constexpr int L_W1 = 0;
constexpr int L_W2 = L_W1 + 32 * S_W1;
constexpr int L_W3 = L_W2 + 64 * S_W2;
constexpr int L_W4 = L_W3 + 32 * S_W3;
constexpr int L_W5 = L_W4 + 64 * S_W4;
constexpr int L_END = L_W5 + 64 * S_W5;  // 10496 floats

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

// Forward of one 16-sample tile starting at sample `s0` (samples >= n are zero).
__device__ __forceinline__ void forward_tile(const float *lds, const float *__restrict__ enc, bool level_major,
                                             const float *__restrict__ dirs, uint32_t n, size_t s0, int i, int g,
                                             Fwd &f) {
    const size_t s = s0 + i;
    const bool live = s < n;
#pragma unroll
    for (int b = 0; b < 2; ++b) f.x[b] = live ? load_enc4(enc, level_major, n, s, b, g) : v4f{0.0f, 0.0f, 0.0f, 0.0f};
    layer_fwd<32, 64, S_W1>(lds + L_W1, f.x, f.h1, i, g);
    relu<4>(f.h1);
    layer_fwd<64, 16, S_W2>(lds + L_W2, f.h1, &f.c[0], i, g);
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    if (live) {
        dx = dirs[3 * s + 0];
        dy = dirs[3 * s + 1];
        dz = dirs[3 * s + 2];
    }
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
    for (uint32_t t = wave; t < n_tiles; t += n_waves) {
        const size_t s0 = (size_t)t * kTile;
        Fwd f;
        forward_tile(lds, enc, level_major, dirs, n, s0, i, g, f);
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

__global__ void __launch_bounds__(kThreads)
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
    const uint32_t wave = blockIdx.x * kWaves + wv;
    const uint32_t n_waves = gridDim.x * kWaves;

    v4f gW1[8], gW2[4], gW3[8], gW4[16], gW5[4];
#pragma unroll
    for (int k = 0; k < 8; ++k) gW1[k] = gW3[k] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) gW2[k] = gW5[k] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 16; ++k) gW4[k] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

    for (uint32_t t = wave; t < n_tiles; t += n_waves) {
        const size_t s0 = (size_t)t * kTile;
        const size_t s = s0 + i;
        const bool live = s < n;
        Fwd f;
        forward_tile(lds, enc, level_major, dirs, n, s0, i, g, f);
        const v4f gin = live ? *reinterpret_cast<const v4f *>(dL_ddrgbs + 4 * s) : v4f{0.0f, 0.0f, 0.0f, 0.0f};
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
        wave_lds_sync();
        dw_accumulate<4, 1>(gW5, zs, ys, i, g);
        v4f d3[4];
        layer_bwd<64, 16, S_W5>(lds + L_W5, &dr, d3, i, g);
        relu_mask<4>(d3, f.h3);
        wave_lds_sync();
        // W4
        tile_to_lds<4>(zs, f.h2, i, g);
        tile_to_lds<4>(ys, d3, i, g);
        wave_lds_sync();
        dw_accumulate<4, 4>(gW4, zs, ys, i, g);
        v4f d2[4];
        layer_bwd<64, 64, S_W4>(lds + L_W4, d3, d2, i, g);
        relu_mask<4>(d2, f.h2);
        wave_lds_sync();
        // W3
        tile_to_lds<2>(zs, f.c, i, g);
        tile_to_lds<4>(ys, d2, i, g);
        wave_lds_sync();
        dw_accumulate<2, 4>(gW3, zs, ys, i, g);
        v4f dc[2];
        layer_bwd<32, 64, S_W3>(lds + L_W3, d2, dc, i, g);
        if (g == 0) dc[0][0] += dsig;
        wave_lds_sync();
        // W2
        tile_to_lds<4>(zs, f.h1, i, g);
        tile_to_lds<1>(ys, &dc[0], i, g);
        wave_lds_sync();
        dw_accumulate<4, 1>(gW2, zs, ys, i, g);
        v4f d1[4];
        layer_bwd<64, 16, S_W2>(lds + L_W2, &dc[0], d1, i, g);
        relu_mask<4>(d1, f.h1);
        wave_lds_sync();
        // W1
        tile_to_lds<2>(zs, f.x, i, g);
        tile_to_lds<4>(ys, d1, i, g);
        wave_lds_sync();
        dw_accumulate<2, 4>(gW1, zs, ys, i, g);
        v4f dx[2];
        layer_bwd<32, 64, S_W1>(lds + L_W1, d1, dx, i, g);
        if (live) {
#pragma unroll
            for (int b = 0; b < 2; ++b) store_enc4(dL_denc, level_major, n, s, b, g, dx[b]);
        }
        wave_lds_sync();
    }

    // block reduction of the weight gradients in LDS (reusing the weight area),
    // then one workspace row per workgroup
    __syncthreads();
    float *red = lds;
    for (int t = threadIdx.x; t < P_END; t += kThreads) red[t] = 0.0f;
    __syncthreads();
    // waves add in a fixed order, so the result does not depend on scheduling
    for (int w = 0; w < kWaves; ++w) {
        if (wv == w) {
            dw_to_lds<2, 4>(red + P_W1, gW1, 64, i, g);
            dw_to_lds<4, 1>(red + P_W2, gW2, 16, i, g);
            dw_to_lds<2, 4>(red + P_W3, gW3, 64, i, g);
            dw_to_lds<4, 4>(red + P_W4, gW4, 64, i, g);
            dw_to_lds<4, 1>(red + P_W5, gW5, 3, i, g);
        }
        __syncthreads();
    }
    float *row = workspace + (size_t)blockIdx.x * P_END;
    for (int t = threadIdx.x; t < P_END; t += kThreads) row[t] = red[t];
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
    if (d.enc_layout > NGP_ENC_LEVEL_MAJOR) {
        set_error(kBadDescriptor, "%s: enc_layout=%u must be 0 or 1", target, d.enc_layout);
        return;
    }
    if (d.n_samples == 0) return;
    if (!check_buffers(buffers, 4, target)) return;
    if (!aligned16(buffers[0]) || !aligned16(buffers[3])) {
        set_error(kBadArgument, "%s: encodings and drgbs must be 16-byte aligned", target);
        return;
    }
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(forward_grid(d.n_samples)), dim3(kThreads), 0, (hipStream_t)stream_,
                       d.n_samples, d.enc_layout == NGP_ENC_LEVEL_MAJOR, (const float *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                       (float *)buffers[3]);
    check_launch(target);
}

extern "C" void nerf_mlp_backward(ngp_stream_t stream_, void **buffers, const char *opaque, size_t opaque_len) {
    const char *target = "nerf_mlp_backward";
    NerfMlpDescriptor d;
    if (!read_descriptor(opaque, opaque_len, &d, target)) return;
    if (d.n_workspace_blocks == 0 || d.enc_layout > NGP_ENC_LEVEL_MAJOR) {
        set_error(kBadDescriptor, "%s: n_workspace_blocks must be positive and enc_layout 0 or 1", target);
        return;
    }
    if (!check_buffer_list(buffers, {2, 5, 6}, target)) return;
    if (d.n_samples > 0 && !check_buffers(buffers, 7, target)) return;
    if (d.n_samples > 0 && (!aligned16(buffers[0]) || !aligned16(buffers[3]) || !aligned16(buffers[4]))) {
        set_error(kBadArgument, "%s: encodings, dL_ddrgbs and dL_dencodings must be 16-byte aligned", target);
        return;
    }
    hipStream_t stream = (hipStream_t)stream_;
    hipLaunchKernelGGL(mlp_backward_kernel, dim3(d.n_workspace_blocks), dim3(kThreads), 0, stream, d.n_samples,
                       d.enc_layout == NGP_ENC_LEVEL_MAJOR, (const float *)buffers[0], (const float *)buffers[1], (const float *)buffers[2],
                       (const float *)buffers[3], (float *)buffers[4], (float *)buffers[6]);
    if (!check_launch(target)) return;
    hipLaunchKernelGGL(sum_rows_kernel, dim3(div_up(P_END, kWave)), dim3(kSumWaves * kWave), 0, stream,
                       d.n_workspace_blocks,
                       (const float *)buffers[6], (float *)buffers[5]);
    check_launch(target);
}
