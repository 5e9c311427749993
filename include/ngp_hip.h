/*
 * ngp_hip.h — C ABI of libngp_hip.so, the gfx950 kernels for the NeRF hot path
 * (multiresolution hash encoder, occupancy-grid ray marcher, volume integration,
 * inference renderer and the fused NeRF MLP).
 *
 * Every compute entry point has the legacy XLA custom-call signature
 *
 *     void target(stream, void **buffers, const char *opaque, size_t opaque_len)
 *
 * `buffers` lists the operands in call order followed by the results in result
 * order; every buffer is device memory owned by the caller.  `opaque` is the raw
 * bytes of the target's descriptor struct below and must be exactly
 * sizeof(descriptor) long.  Targets never throw: on a bad descriptor, a null
 * buffer or a failed launch they record an error that ngp_get_last_error()
 * returns (and clears), and launch nothing.  The error state is process-wide, so
 * a thread other than the one that ran the target (e.g. the host thread of a
 * framework whose executor threads run the targets) sees it.
 *
 * Work is enqueued asynchronously on `stream` (a hipStream_t; NULL is the legacy
 * default stream).  No target synchronises the stream.
 */
#ifndef NGP_HIP_H
#define NGP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A hipStream_t, passed as an opaque pointer so this header needs no HIP headers. */
typedef void *ngp_stream_t;

/* ---------------------------------------------------------------------------
 * Descriptors.  Field order, types and sizes follow the reference's
 * deps/volume-rendering-jax/lib/impl/volrend.h:23-118 so that the opaque bytes a
 * reference-side binding builds are accepted unchanged.
 * ------------------------------------------------------------------------- */

/* volrend.h:23-29 */
typedef struct IntegratingDescriptor {
    uint32_t n_rays;
    uint32_t total_samples;
} IntegratingDescriptor;

/* volrend.h:32-42 */
typedef struct IntegratingBackwardDescriptor {
    uint32_t n_rays;
    uint32_t total_samples;
    float near_distance;
} IntegratingBackwardDescriptor;

/* volrend.h:45-54 */
typedef struct IntegratingInferenceDescriptor {
    uint32_t n_total_rays;
    uint32_t n_rays;
    uint32_t march_steps_cap;
} IntegratingInferenceDescriptor;

/* volrend.h:57-86 */
typedef struct MarchingDescriptor {
    uint32_t n_rays;
    uint32_t total_samples;
    uint32_t diagonal_n_steps;
    uint32_t K;
    uint32_t G;
    float bound;
    float stepsize_portion;
} MarchingDescriptor;

/* volrend.h:89-108 */
typedef struct MarchingInferenceDescriptor {
    uint32_t n_total_rays;
    uint32_t n_rays;
    uint32_t diagonal_n_steps;
    uint32_t K;
    uint32_t G;
    uint32_t march_steps_cap;
    float bound;
    float stepsize_portion;
} MarchingInferenceDescriptor;

/* volrend.h:110-113 */
typedef struct Morton3DDescriptor {
    uint32_t length;
} Morton3DDescriptor;

/* volrend.h:116-118 */
typedef struct PackbitsDescriptor {
    uint32_t n_bytes;
} PackbitsDescriptor;

/* Hash grid (no native counterpart in the reference: models/encoders.py:57-256 is
 * pure JAX).  Level scales are computed on the host in float64 and rounded to f32
 * (encoders.py:91, 216); resolutions and row offsets come from the same host loop
 * (encoders.py:89-103).  They are never recomputed on the device. */
#define NGP_HASHGRID_MAX_LEVELS 32
#define NGP_DTYPE_F32 0u
#define NGP_DTYPE_F16 1u
/* Layout of encodings and their gradients: point-major [N, L*F] (the reference's
 * HashGridEncoder output, enc[n, l*F + f]) or level-major [L, N, F]
 * (enc[(l*N + n)*F + f]; used inside the fused training step, where it keeps the
 * per-level work of the encoder on one XCD with coalesced accesses). */
#define NGP_ENC_POINT_MAJOR 0u
#define NGP_ENC_LEVEL_MAJOR 1u
typedef struct HashGridDescriptor {
    uint32_t n_points;          /* N */
    uint32_t L;                 /* levels, 1..NGP_HASHGRID_MAX_LEVELS */
    uint32_t F;                 /* features per entry: 1, 2, 4 or 8 */
    uint32_t T;                 /* max entries per level */
    uint32_t first_hash_level;  /* levels [0, first_hash_level) are dense */
    uint32_t table_dtype;       /* NGP_DTYPE_F32 or NGP_DTYPE_F16 */
    uint32_t n_rows;            /* offsets[L], rows of the table */
    float bound;                /* positions are normalised as (pos + bound) / (2 bound) */
    uint32_t enc_layout;        /* NGP_ENC_POINT_MAJOR or NGP_ENC_LEVEL_MAJOR */
    float scales[NGP_HASHGRID_MAX_LEVELS];
    uint32_t resolutions[NGP_HASHGRID_MAX_LEVELS];
    uint32_t offsets[NGP_HASHGRID_MAX_LEVELS + 1];
} HashGridDescriptor;

/* jax-tcnn's hash-grid encoding (the reference's inference-time encoder,
 * models/encoders.py:259-305, bound through deps/jax-tcnn/lib/ffi.cc and
 * lib/impl/hashgrid.cu, which run tiny-cuda-nn v1.6's grid kernels).  The
 * descriptor is jaxtcnn::HashGridDescriptor byte for byte
 * (deps/jax-tcnn/lib/impl/tcnnutils.h:11-29).  Levels are unpadded: offsets[l+1] -
 * offsets[l] = res_l^3 for dense levels, T for hashed ones. */
typedef struct TcnnHashGridDescriptor {
    uint32_t n_coords;      /* N */
    uint32_t L;             /* levels (1..32) */
    uint32_t F;             /* features per level: 2 or 4 */
    uint32_t N_min;         /* base resolution */
    float per_level_scale;  /* b */
} TcnnHashGridDescriptor;

/* Replaces jaxtcnn::hashgrid_encode (hashgrid.cu:20-96; operands and results as
 * hashgrid.cu:35-42).
 * in : offset_table u32[L+1] (device), coords f32[3, N], params f32[offsets[L], F]
 * out: encodings f32[L*F, N], dy_dcoords f32[L*F, N, 3] */
void tcnn_hashgrid_encode(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Replaces jaxtcnn::hashgrid_encode_backward (hashgrid.cu:98-175; buffers as
 * hashgrid.cu:111-118).
 * in : offset_table u32[L+1], coords f32[3, N], dL_dy f32[L*F, N], dy_dcoords f32[L*F, N, 3]
 * out: dL_dparams f32[offsets[L], F] (fully written), dL_dcoords f32[3, N] */
void tcnn_hashgrid_encode_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Sky-sphere background model of unbounded scenes (models/nerfs.py:133-213, built by
 * make_skysphere_background_model_ngp, nerfs.py:383-419): the ray's far intersection
 * with the sphere |x| = radius (radius = 4 bound), SH degree 2 of the hit point
 * divided by the Frobenius norm of ALL the batch's hit points (nerfs.py:191-196:
 * jnp.linalg.norm over the whole [n, 3] array, + 1e-15), SH degree 4 of the ray
 * direction and the appearance embedding feed a 32->32->3 MLP without biases (ReLU
 * hidden layers, sigmoid output); rays whose intersection t is not finite get 0.
 * A ray that misses the sphere makes that norm NaN and with it every other ray's
 * colour, as in the reference.  The parameters are one flat f32 array of the flax
 * Dense kernels ([in, out], row-major): rgb_mlp/Dense_0 [20+E, 32],
 * Dense_1 [32, 32], Dense_2 [32, 3]. */
#define NGP_SKY_MAX_EXTRA_DIMS 16u
typedef struct SkySphereBgDescriptor {
    uint32_t n_rays;
    uint32_t n_extra_dims;       /* E, appearance embedding width, 0..NGP_SKY_MAX_EXTRA_DIMS */
    uint32_t appearance_per_ray; /* 1: appearance f32[n_rays, E]; 0: one f32[E] for every ray */
    float radius;
} SkySphereBgDescriptor;

/* Number of f32 parameters: (20 + E) * 32 + 32 * 32 + 32 * 3. */
uint32_t sky_sphere_bg_n_params(uint32_t n_extra_dims);

/* Bytes of the workspace both sky-sphere targets take (256-byte aligned device
 * memory); 0 with an error recorded for an invalid descriptor. */
size_t sky_sphere_bg_workspace_size(const char *opaque, size_t opaque_len);

/* Replaces SkySphereBg.__call__ (nerfs.py:160-213).  opaque: SkySphereBgDescriptor.
 * in : rays_o f32[n, 3], rays_d f32[n, 3], appearance (see appearance_per_ray; may
 *      be NULL when E = 0), params f32[sky_sphere_bg_n_params(E)]
 * out: colors f32[n, 3], workspace u8[sky_sphere_bg_workspace_size(opaque)] */
void sky_sphere_bg(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* The VJP of sky_sphere_bg with respect to the parameters and the per-ray
 * appearance embeddings (the gradient jax.grad takes through bg_fn in
 * app/nerf/_utils.py:128-133); rays and the batch norm get none.  Weight
 * gradients are summed in a fixed order (deterministic).  Needs appearance_per_ray
 * = 1 when E > 0.
 * in : rays_o, rays_d, appearance, params (as sky_sphere_bg), dL_dcolors f32[n, 3]
 * out: dL_dparams f32[n_params], dL_dappearance f32[n, E] (may be NULL when E = 0),
 *      workspace u8[sky_sphere_bg_workspace_size(opaque)] */
void sky_sphere_bg_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Fused NeRF MLP of models/nerfs.py:39-86 with make_nerf_ngp's shapes
 * (nerfs.py:422-454): density MLP 32->64->16, SH degree 4 of the direction,
 * colour MLP (16+16)->64->64->3, no biases, ReLU hidden layers, trunc_exp density
 * and sigmoid colour.  The parameters are one flat f32 array laid out as the flax
 * Dense kernels ([in, out], row-major) in this order:
 *   density_mlp/Dense_0 [32,64], density_mlp/Dense_1 [64,16],
 *   rgb_mlp/Dense_0 [32,64], rgb_mlp/Dense_1 [64,64], rgb_mlp/Dense_2 [64,3]. */
#define NGP_MLP_N_PARAMS 9408u
typedef struct NerfMlpDescriptor {
    uint32_t n_samples;
    uint32_t n_workspace_blocks;  /* rows of the backward workspace (0 for forward) */
    uint32_t enc_layout;          /* layout of encodings / dL_dencodings (L=16, F=2) */
    uint32_t param_dtype;         /* NGP_DTYPE_F32, or NGP_DTYPE_F16: f16 weights and f16
                                     MFMA operands, f32 accumulation and activations */
} NerfMlpDescriptor;

/* ---------------------------------------------------------------------------
 * Entry points replacing the reference's custom-call targets
 * (deps/volume-rendering-jax/lib/ffi.cc:25-51, volrend.h:121-175).
 * ------------------------------------------------------------------------- */

/* marching.cu:435-522.  opaque: MarchingDescriptor.
 * in : rays_o f32[n,3], rays_d f32[n,3], t_starts f32[n], t_ends f32[n],
 *      noises f32[n], occupancy_bitfield u8[K*G^3/8]
 * out: next_sample_write_location u32[1], number_of_exceeded_samples u32[1],
 *      ray_is_valid bool[n], rays_n_samples u32[n], rays_sample_startidx u32[n],
 *      idcs u32[S], xyzs f32[S,3], dirs f32[S,3], dss f32[S], z_vals f32[S]
 * Sample ranges are assigned by an exclusive prefix sum in ray order (the
 * reference assigns them with atomicAdd in scheduling order). */
void march_rays(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* marching.cu:524-604.  opaque: MarchingInferenceDescriptor.
 * in : rays_o f32[N,3], rays_d f32[N,3], t_starts f32[N], t_ends f32[N],
 *      occupancy_bitfield u8[K*G^3/8], next_ray_index_in u32[1],
 *      terminated bool[n], indices_in u32[n]
 * out: next_ray_index u32[1], indices_out u32[n], n_samples u32[n],
 *      t_starts_out f32[n], xyzs f32[n,cap,3], dss f32[n,cap], z_vals f32[n,cap]
 * Terminated slots receive new rays in slot order.  indices_out must not alias
 * indices_in (next_ray_index may alias next_ray_index_in). */
void march_rays_inference(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* marching.cu:606-636.  opaque: Morton3DDescriptor.  in: xyzs u32[len,3]; out: idcs u32[len]. */
void morton3d(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* marching.cu:638-665.  opaque: Morton3DDescriptor.  in: idcs u32[len]; out: xyzs u32[len,3]. */
void morton3d_invert(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* packbits.cu:36-72.  opaque: PackbitsDescriptor.
 * in : density_threshold f32[8*n_bytes], density_grid f32[8*n_bytes]
 * out: occupied_mask bool[8*n_bytes], occupancy_bitfield u8[n_bytes] */
void pack_density_into_bits(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* integrating.cu:325-377.  opaque: IntegratingDescriptor.
 * in : rays_sample_startidx u32[n], rays_n_samples u32[n], bgs f32[n,3],
 *      dss f32[S], z_vals f32[S], drgbs f32[S,4]
 * out: measured_batch_size u32[1], final_rgbds f32[n,4], final_opacities f32[n] */
void integrate_rays(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* integrating.cu:379-446.  opaque: IntegratingBackwardDescriptor.
 * in : rays_sample_startidx u32[n], rays_n_samples u32[n], bgs f32[n,3],
 *      dss f32[S], z_vals f32[S], drgbs f32[S,4], final_rgbds f32[n,4],
 *      final_opacities f32[n], dL_dfinal_rgbds f32[n,4]
 * out: dL_dbgs f32[n,3], dL_dz_vals f32[S], dL_ddrgbs f32[S,4] */
void integrate_rays_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* integrating.cu:448-507.  opaque: IntegratingInferenceDescriptor.
 * in : rays_bg f32[N,3], rays_rgbd f32[N,4], rays_T f32[N], n_samples u32[n],
 *      indices u32[n], dss f32[n,cap], z_vals f32[n,cap], drgbs f32[n,cap,4]
 * out: terminate_cnt u32[1], terminated bool[n], rays_rgbd_out f32[n,4],
 *      rays_T_out f32[n] */
void integrate_rays_inference(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* ---------------------------------------------------------------------------
 * Entry points with no native counterpart in the reference (its hash encoder and
 * MLPs are XLA-generated code).  Same calling convention.
 * ------------------------------------------------------------------------- */

/* Forward of HashGridEncoder.__call__ (models/encoders.py:82-233).
 * opaque: HashGridDescriptor.
 * in : pos f32[N,3], table f32|f16[n_rows,F]
 * out: encodings f32[N, L*F] (enc_layout NGP_ENC_POINT_MAJOR) or f32[L, N, F]
 *      (NGP_ENC_LEVEL_MAJOR) */
void hashgrid_encode(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Table gradient of HashGridEncoder (the VJP of the gather at encoders.py:226 and
 * of the weighted sum at encoders.py:231).  opaque: HashGridDescriptor.
 * in : pos f32[N,3], dL_dencodings f32 in the descriptor's enc_layout
 * out: dL_dtable f32[n_rows, F]   (overwritten, not accumulated into; 16-byte aligned),
 *      workspace u8[hashgrid_encode_backward_workspace_size(opaque)], 256-byte
 *      aligned (unused, may be NULL, when N == 0)
 * The gradient is scattered without global atomics: contributions are sorted
 * into row slices that one workgroup sums in LDS. */
void hashgrid_encode_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Bytes of the workspace hashgrid_encode_backward needs for the descriptor
 * `opaque` (no counterpart in the reference, whose scatter uses atomics on the
 * output).  Returns 0 and sets the last error for an invalid descriptor. */
size_t hashgrid_encode_backward_workspace_size(const char *opaque, size_t opaque_len);

/* Forward of NeRF.__call__ after the position encoder (models/nerfs.py:66-86).
 * opaque: NerfMlpDescriptor.
 * in : encodings f32[S,32] or f32[16,S,2] (enc_layout), dirs f32[S,3],
 *      params f32 or f16 [NGP_MLP_N_PARAMS] (param_dtype)
 * out: drgbs f32[S,4]   (density, r, g, b) */
void nerf_mlp_forward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Backward of nerf_mlp_forward (JAX autodiff of nerfs.py:66-86 with trunc_exp's
 * custom VJP, nerfs.py:221-238).  opaque: NerfMlpDescriptor.
 * in : encodings (as forward), dirs f32[S,3], params (as forward),
 *      dL_ddrgbs f32[S,4] (loss-scaled on the f16 path)
 * out: dL_dencodings (same layout as encodings), dL_dparams f32[NGP_MLP_N_PARAMS],
 *      workspace f32[n_workspace_blocks, NGP_MLP_N_PARAMS] */
void nerf_mlp_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* No counterpart in the reference: nerf_mlp_backward for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same descriptor, buffers and results as nerf_mlp_backward
 * without the workspace buffer:
 * in : encodings, dirs, params, dL_ddrgbs;  out: dL_dencodings, dL_dparams.
 * The workspace is a fixed device array of NGP_MLP_NO_ALLOC_MAX_BLOCKS rows of
 * NGP_MLP_N_PARAMS floats (18.4 MiB per GPU, held until the process ends), so
 * n_workspace_blocks must be at most NGP_MLP_NO_ALLOC_MAX_BLOCKS (else the call
 * records an error and launches nothing).  Nothing is allocated or freed by the
 * call.  Every call on a GPU shares that GPU's array: calls that overlap on one
 * GPU (on different streams, or from several threads without ordering) race. */
#define NGP_MLP_NO_ALLOC_MAX_BLOCKS 512u
void nerf_mlp_backward_no_alloc(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct NerfFieldChunksDescriptor {
    HashGridDescriptor grid;  /* the NeRF position encoder (L = 16, F = 2); n_points and
                                 enc_layout are not used */
    uint32_t n_slots;         /* ray-pool slots */
    uint32_t cap;             /* march_steps_cap: samples per slot */
    uint32_t n_total_rays;    /* rows of rays_d */
    uint32_t param_dtype;     /* MLP weights: NGP_DTYPE_F32 or NGP_DTYPE_F16 */
} NerfFieldChunksDescriptor;

/* The field (hash encoding + NeRF MLP, models/nerfs.py:66-86) on one round of the
 * inference ray pool, fused: slot j holds n_samples[j] samples of ray indices[j]
 * (the outputs of march_rays_inference); padding entries get drgbs = 0.  Equal to
 * hashgrid_encode followed by nerf_mlp_forward on the marched samples.  No
 * counterpart in the reference, which evaluates its field on every [slot, step]
 * entry, padding included (models/renderers/cuda.py:213-223).
 * in : xyzs f32[n_slots, cap, 3], n_samples u32[n_slots], indices u32[n_slots],
 *      rays_d f32[n_total_rays, 3], table f32|f16[offsets[16], 2],
 *      params f32|f16[NGP_MLP_N_PARAMS]
 * out: drgbs f32[n_slots, cap, 4] */
void nerf_field_chunks(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct OccupancyGridUpdateDescriptor {
    HashGridDescriptor grid; /* the NeRF position encoder (L = 16, F = 2); n_points and
                                enc_layout are not used, grid.bound is the scene bound */
    uint32_t G;              /* grid resolution (power of two, <= 1024) */
    uint32_t K;              /* cascades: density f32[K G^3], bitfield u8[K G^3 / 8] */
    uint32_t cascade;        /* the cascade this call updates */
    uint32_t n_alive;        /* alive cells of the cascade */
    uint32_t update_all;     /* 1: every alive cell once; 0: 2 * max(1, max(1, n_alive / 2) / 2)
                                random cells, the first half uniform over the alive cells,
                                the second half uniform over the occupied ones */
    uint32_t param_dtype;    /* MLP weights: NGP_DTYPE_F32 or NGP_DTYPE_F16 */
    uint32_t seed_lo, seed_hi, step; /* the draws' key */
} OccupancyGridUpdateDescriptor;

/* NeRFState.update_ogrid_density (utils/types.py:1149-1225) for one cascade: the
 * alive cells' densities are multiplied by 0.95, then each chosen cell's density
 * becomes max(its decayed density, the field's density at a random point of the
 * cell).  The point of cell m (Morton index within the cascade) is
 * (morton3d_invert(m) / (G - 1) * 2 - 1) * (mip - mip / G) + (2 u - 1) mip / G per
 * axis, mip = min(bound, 2^cascade), u uniform in [0, 1) (types.py:1193-1207); the
 * density is trunc_exp of the density MLP's first output on the position encoding
 * (models/nerfs.py:66-77; only 32 -> 64 -> 16 runs).  A cell chosen several times
 * keeps the largest of its densities (the reference's scatter keeps one of them,
 * order unspecified).  The occupied cells are those set in the bitfield (the last
 * threshold; they are alive by construction); with none, the second half is
 * uniform over the alive cells as well.  Random numbers: r_s = splitmix64 draw of
 * (seed, step, key = cascade << 32 | slot, stream s); a uniform choice among n is
 * (r >> 32) * n >> 32; the first half (and an empty occupied set) chooses alive
 * cells with r_0, the second half occupied cells with r_1 (in Morton order), and
 * the point's u for x, y, z are bits 43-63, 22-42 and 1-21 of r_2 over 2^21.
 * in : alive u32[n_alive] (Morton indices within the cascade, ascending),
 *      occupancy_bitfield u8[K G^3 / 8], table f32|f16[offsets[16], 2],
 *      params f32|f16[NGP_MLP_N_PARAMS]
 * inout: density f32[K G^3]
 * out: scratch u32[G^3 / 64 + 1] (occupied-cell prefix counts; not used, and may
 *      be NULL, when update_all) */
void occupancy_grid_update(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct NgpPoolScatterDescriptor {
    uint32_t n;       /* pool slots */
    uint32_t n_total; /* rows of dst */
    uint32_t width;   /* floats per row */
} NgpPoolScatterDescriptor;

/* dst[idx[i]] = src[i] for every slot whose idx[i] < n_total (the wrappers'
 * `.at[indices].set(...)` with out-of-range indices dropped,
 * volrendjax/marching/__init__.py:139-157, integrating/__init__.py:97-110).
 * in : idx u32[n], src f32[n, width]   inout: dst f32[n_total, width] */
void pool_scatter(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* ---------------------------------------------------------------------------
 * Learned proposal (no counterpart in the reference): a hash grid (L = 5, F = 2)
 * and a 10 -> 16 -> 1 MLP predicting density at 64 equal bins of each ray's
 * segment [t_start, t_end), trained against the NeRF's sample weights.
 * ------------------------------------------------------------------------- */

#define NGP_PROPOSAL_BINS 64u
#define NGP_PROPOSAL_N_PARAMS 176u /* W1 [10][16], W2 [16][1], no biases */

typedef struct ProposalRaysDescriptor {
    uint32_t n_rays;
} ProposalRaysDescriptor;

/* Bin centres o + (t0 + (b + 1/2) (t1 - t0) / 64) d.  opaque: ProposalRaysDescriptor.
 * in : rays_o f32[n,3], rays_d f32[n,3], t_starts f32[n], t_ends f32[n]
 * out: points f32[n, 64, 3] */
void proposal_points(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct ProposalMlpDescriptor {
    uint32_t n_points;
    uint32_t n_workspace_blocks; /* rows of the backward workspace (0 for forward) */
    uint32_t enc_layout;         /* encodings [N, 10] or level-major [5, N, 2] */
    uint32_t param_dtype;        /* NGP_DTYPE_F32 or NGP_DTYPE_F16 weights */
} ProposalMlpDescriptor;

/* sigma = exp(relu(x W1) W2).  opaque: ProposalMlpDescriptor.
 * in : encodings f32, params f32|f16[NGP_PROPOSAL_N_PARAMS]   out: sigma f32[N] */
void proposal_mlp_forward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Backward of proposal_mlp_forward (exp's VJP clipped at |logit| <= 15).
 * in : encodings, params, dL_dsigma f32[N]
 * out: dL_dencodings (layout of encodings), dL_dparams f32[NGP_PROPOSAL_N_PARAMS],
 *      workspace f32[n_workspace_blocks, NGP_PROPOSAL_N_PARAMS] */
void proposal_mlp_backward(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct ProposalLossDescriptor {
    uint32_t n_rays;
    float eps; /* 1e-7 */
} ProposalLossDescriptor;

/* Histogram-bound loss.  Per valid ray the NeRF's sample weights w_i = T_i alpha_i
 * (0 once T falls to 1e-4) are summed into 64 equal bins of [t0, t1) (W_b); the
 * proposal's bin weights are w_b = T_b (1 - exp(-sigma_b D)), D = (t1 - t0) / 64;
 * loss = sum over valid rays and bins of max(0, W_b - w_b)^2 / (W_b + eps) / n_valid.
 * in : rays_sample_startidx u32[n], rays_n_samples u32[n], dss f32[S], z_vals f32[S],
 *      drgbs f32[S,4] (the NeRF's), t_starts f32[n], t_ends f32[n],
 *      sigma_p f32[n, 64], ray_is_valid bool[n], n_valid_rays u32[1]
 * out: dL_dsigma_p f32[n, 64], loss f32[1] */
void proposal_loss(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

#define NGP_PROPOSAL_QUANTILES 0u /* samples at the quantiles of the bin weights */
#define NGP_PROPOSAL_SUPPORT 1u   /* samples evenly over the bins that hold the mass */

typedef struct ProposalSampleDescriptor {
    uint32_t n_rays;
    uint32_t n_per_ray;   /* samples per ray (48), at most 64 */
    float mass_threshold; /* below it the ray gets no samples and is marked (0.98 renders) */
    uint32_t mode;        /* NGP_PROPOSAL_QUANTILES or NGP_PROPOSAL_SUPPORT */
    float eps;            /* SUPPORT: mass share left out at either end (1e-3) */
    /* marching statics of the occupancy grid (as MarchingDescriptor) */
    uint32_t diagonal_n_steps, K, G;
    float bound, stepsize_portion;
    float empty_mass;     /* a ray with M < empty_mass gets no samples and no fallback: it
                             renders as background (0: off, every ray below
                             mass_threshold falls back) */
} ProposalSampleDescriptor;

/* Samples placed by the proposal: per ray, the bin weights w_b of sigma_p give the
 * mass M = sum w_b; a ray with t_end > t_start and M >= mass_threshold gets n_per_ray
 * samples at the (j + 1/2) / n_per_ray quantiles of the piecewise constant density
 * w_b / (M D), ds_j = t_{j+1} - t_j (the last: min(D / (n p_b), t_end - t_j)), or
 * (SUPPORT) evenly over the bins from the first whose cumulative share exceeds eps
 * to the first where it reaches 1 - eps, ds = the spacing.  Samples whose cell the
 * occupancy grid marks empty are dropped (march_rays never samples such cells): the
 * others are packed in order at the front of the ray's slots, rays_n_samples of
 * them, the remaining slots are zero.  A ray
 * below the threshold gets none and fallback = true (the occupancy-grid marcher
 * renders it), unless its mass is below empty_mass: then fallback = false and the ray
 * composites to its background.  Sample arrays use [n_rays, n_per_ray] slots, startidx = r n_per_ray.
 * in : rays_o f32[n,3], rays_d f32[n,3], t_starts f32[n], t_ends f32[n], sigma_p f32[n,64],
 *      occupancy_bitfield u8[K*G^3/8]
 * out: rays_sample_startidx u32[n], rays_n_samples u32[n], fallback bool[n],
 *      xyzs f32[n,N,3], dss f32[n,N], z_vals f32[n,N] */
void proposal_sample(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct ProposalRenderDescriptor {
    HashGridDescriptor grid;          /* the proposal's encoder (L = 5, F = 2) */
    ProposalSampleDescriptor sample;  /* placement, thresholds, marching statics */
    uint32_t param_dtype;             /* proposal MLP weights: f32 or f16 */
} ProposalRenderDescriptor;

/* proposal_points + hashgrid_encode + proposal_mlp_forward + proposal_sample fused
 * for rendering: one wavefront per ray, lane b evaluates the proposal at bin b's
 * centre, then the samples are placed; nothing in between goes through memory.
 * in : rays_o f32[n,3], rays_d f32[n,3], t_starts f32[n], t_ends f32[n],
 *      table f32|f16[offsets[5], 2], params f32|f16[NGP_PROPOSAL_N_PARAMS],
 *      occupancy_bitfield u8[K*G^3/8]
 * out: as proposal_sample */
void proposal_render_samples(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct ProposalTrainDescriptor {
    ProposalRenderDescriptor render; /* the proposal, quantile placement, the mass threshold
                                        (0 in training: every ray gets its samples) and
                                        the marching statics; render.sample.n_rays rays */
    uint32_t total_samples;          /* S, as march_rays */
    uint32_t n_uniform;              /* U: samples added per ray, stratified uniformly
                                        over [t_start, t_end); n_per_ray + U <= 64 */
} ProposalTrainDescriptor;

/* Training samples placed by the proposal instead of the occupancy-grid marcher
 * (no counterpart in the reference), in march_rays' buffer layout so the rest of
 * the step runs unchanged: per ray, the proposal's 64 bin densities give n_per_ray
 * samples at the (j + noise) / n_per_ray quantiles of its predicted mass
 * (stratified), plus n_uniform samples at t_start + (i + noise) (t_end - t_start) / U
 * (so the field also sees where the proposal predicts no mass), sorted, each with
 * ds = the distance to the next sample (the last: to t_end, at most the uniform
 * spacing); those in cells the occupancy grid marks empty are dropped, and the
 * rays' kept samples are packed in ray order; a ray is valid iff t_end > t_start
 * and its samples fit the S slots after those of the rays before it.
 * next_sample_write_location = the samples every ray asked for,
 * number_of_exceeded_samples = those of the rays that did not fit.
 * in : rays_o f32[n,3], rays_d f32[n,3], t_starts f32[n], t_ends f32[n],
 *      noises f32[n], table f32|f16[offsets[5], 2], params f32|f16[NGP_PROPOSAL_N_PARAMS],
 *      occupancy_bitfield u8[K*G^3/8]
 * out: as march_rays (next_sample_write_location u32[1], number_of_exceeded_samples
 *      u32[1], ray_is_valid bool[n], rays_n_samples u32[n], rays_sample_startidx
 *      u32[n], idcs u32[S], xyzs f32[S,3], dirs f32[S,3], dss f32[S], z_vals f32[S]),
 *      then workspace f32[5, n, 64] (scratch: the counting pass stages each kept
 *      sample there, so the proposal is evaluated once per ray; contents undefined
 *      afterwards; may be omitted when n = 0) */
void proposal_train_samples(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* ---------------------------------------------------------------------------
 * Training-step harness (the device-side counterpart of app/nerf/_utils.py:80-170
 * and app/nerf/train.py:47-48).  Same calling convention.
 * ------------------------------------------------------------------------- */

#define NGP_TRAIN_PERTURB 1u
#define NGP_TRAIN_RANDOM_BG 2u
typedef struct NgpTrainRaysDescriptor {
    uint32_t n_rays;
    uint32_t n_views;
    uint32_t width;
    uint32_t height;
    float fx, fy, cx, cy; /* pinhole intrinsics, pixel centres at +0.5 */
    float bound;          /* scene AABB [-bound, bound]^3 */
    uint32_t seed_lo, seed_hi, step;
    uint32_t flags;       /* NGP_TRAIN_PERTURB | NGP_TRAIN_RANDOM_BG */
    float bg[3];          /* background when NGP_TRAIN_RANDOM_BG is clear */
} NgpTrainRaysDescriptor;

/* Samples n_rays pixels uniformly over all views (with replacement) and builds
 * their rays (utils/types.py:400-439, app/nerf/_utils.py:97-115), AABB near/far
 * (models/renderers/cuda.py:56-97), marching noise U[0,1), background colours and
 * alpha-blended ground truth (utils/data.py:443-464).
 * in : transforms f32[V,12] (camera-to-world rotation row-major, then translation),
 *      rgba u8[V,H,W,4]
 * out: rays_o f32[n,3], rays_d f32[n,3], t_starts f32[n], t_ends f32[n],
 *      noises f32[n], bgs f32[n,3], gt_rgb f32[n,3], pixel_ids u32[n] */
void train_make_rays(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct NgpTrainBgDescriptor {
    uint32_t n_rays;
} NgpTrainBgDescriptor;

/* With a background model (app/nerf/_utils.py:128-133, 150): the ground truth of
 * the pixels train_make_rays sampled, blended over per-ray backgrounds,
 * gt = rgb * alpha + bg * (1 - alpha) (utils/data.py:443-464).
 * in : rgba u8[V,H,W,4], pixel_ids u32[n] (train_make_rays' output), bgs f32[n,3]
 * out: gt_rgb f32[n,3] */
void train_gt_blend(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* The gradient of the loss with respect to those backgrounds: through the
 * composited colour (integrate_rays_backward's dL_dbgs) and through the blended
 * ground truth (-dL_dfinal_rgbds[:3] * (1 - alpha)).
 * in : rgba u8[V,H,W,4], pixel_ids u32[n], dL_dfinal_rgbds f32[n,4], dL_dbgs f32[n,3]
 * out: dL_dbg f32[n,3] */
void train_bg_grad(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* Dynamic loss scaling state, kept in device memory so that no step waits for
 * the host (no counterpart in the reference, which trains in f32 only). */
typedef struct NgpLossScaleState {
    float scale;         /* current loss scale (the f16 path starts at 2^15) */
    uint32_t overflow;   /* set by loss_scale_check when a gradient is not finite */
    uint32_t good_steps; /* finite steps since the scale last changed */
    uint32_t reserved;
} NgpLossScaleState;

#define NGP_LOSS_SCALED 1u /* flags: gradients are multiplied by / divided by state.scale */

typedef struct NgpLossDescriptor {
    uint32_t n_rays;
    float delta;    /* Huber delta (0.1 in app/nerf/_utils.py:154) */
    uint32_t flags; /* NGP_LOSS_SCALED: dL_dfinal_rgbds is multiplied by state.scale */
} NgpLossDescriptor;

/* Huber loss averaged over rgb and over valid rays, and its gradient with respect
 * to final_rgbds (app/nerf/_utils.py:149-158; the depth channel gets 0).
 * in : final_rgbds f32[n,4], gt_rgb f32[n,3], ray_is_valid bool[n]
 * out: n_valid_rays u32[1], dL_dfinal_rgbds f32[n,4], loss f32[1] (unscaled)
 * with NGP_LOSS_SCALED, one more input: NgpLossScaleState[1] */
void train_loss_grad(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct NgpLossScaleCheckDescriptor {
    uint64_t n;
} NgpLossScaleCheckDescriptor;

/* Sets state.overflow if any of the n gradient values is not finite (no
 * counterpart in the reference).
 * in : grads f32[n]   inout: NgpLossScaleState[1] */
void loss_scale_check(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct NgpLossScaleUpdateDescriptor {
    uint32_t growth_interval; /* finite steps before the scale doubles */
    float min_scale, max_scale;
} NgpLossScaleUpdateDescriptor;

/* End of a loss-scaled step: after an overflow the scale halves (not below
 * min_scale) and the count restarts; otherwise the count grows and, at
 * growth_interval, the scale doubles (not above max_scale).  Clears overflow.
 * inout: NgpLossScaleState[1] */
void loss_scale_update(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct NgpAdamDescriptor {
    uint64_t n;
    float lr, b1, b2, eps, eps_root;
    float b1_pow_t, b2_pow_t; /* b1^t and b2^t for the bias corrections, t = update count */
    float weight_decay;       /* added to the (already lr-scaled) update: u += wd * p */
    float grad_scale;         /* gradients are multiplied by this first (loss-scale undo) */
    uint32_t mirror_dtype;    /* NGP_DTYPE_F16: also write the updated params as f16 */
    uint32_t flags;           /* NGP_LOSS_SCALED: gradients are also divided by state.scale,
                                 and nothing is updated when state.overflow is set */
    uint32_t reserved;
} NgpAdamDescriptor;

/* One optax.adam(lr, b1, b2, eps, eps_root) step followed by
 * add_decayed_weights(weight_decay), in that order (app/nerf/_utils.py:19-77).
 * in/out: params f32[n]; in: grads f32[n]; in/out: m f32[n], v f32[n];
 * out: mirror f16[n] (only read when mirror_dtype is NGP_DTYPE_F16; may be NULL);
 * with NGP_LOSS_SCALED, one more input after the mirror slot: NgpLossScaleState[1] */
void adam_update(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

typedef struct HashGridAdamDescriptor {
    HashGridDescriptor grid; /* as hashgrid_encode_backward (F <= 2, f32 table) */
    NgpAdamDescriptor adam;  /* as adam_update; n = n_rows * F, no loss scaling, no mirror */
} HashGridAdamDescriptor;

/* hashgrid_encode_backward followed by adam_update on the table, in one sequence of
 * launches (no counterpart in the reference, whose optimizer step is XLA code): the
 * table gradient's last pass takes each row's optimizer step in place of storing
 * the gradient, so the gradient never goes through memory.  Same results as the two
 * targets one after the other (both use one element update).  Needs the one-pass
 * table gradient (hashgrid_encode_backward_adam_supported returns 1).
 * in : positions f32[N,3], dL_dencodings f32[.., as hashgrid_encode_backward];
 * in/out: params f32[n_rows, F], m f32[n_rows, F], v f32[n_rows, F];
 * scratch: workspace u8[hashgrid_encode_backward_workspace_size(grid)] (256-byte
 * aligned; params, m, v 16-byte aligned).  With N = 0 only the first five are read. */
void hashgrid_encode_backward_adam(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len);

/* 1 if hashgrid_encode_backward_adam takes this HashGridDescriptor (F <= 2, f32
 * table, every level within the one-pass table gradient's slices), else 0. */
int hashgrid_encode_backward_adam_supported(const char *opaque, size_t opaque_len);

/* ---------------------------------------------------------------------------
 * Status.
 * ------------------------------------------------------------------------- */

/* Returns 0 when no error was recorded (by any thread) since the last call,
 * otherwise the code of the first error recorded since then; copies its
 * NUL-terminated message into buf (if buf is non-NULL) and clears the error. */
int ngp_get_last_error(char *buf, size_t buf_len);

/* Library version string. */
const char *ngp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NGP_HIP_H */
