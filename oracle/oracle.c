/*
 * oracle.c — CPU restatement of the reference's NeRF hot-path kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker / the timed
 * CPU baseline.  The product (jaxngp_amd) never calls it.
 *
 * Arithmetic convention (fixed here, followed bit for bit by the HIP kernels):
 *   - float32 throughout, every operation rounded on its own: this file is built
 *     with -ffp-contract=off, and a fused multiply-add appears only where fmaf()
 *     is written out;
 *   - IEEE division;
 *   - exp is ngp_expf() below (Cody-Waite reduction + degree-7 Taylor polynomial),
 *     not the C library's expf, so the CPU and the GPU round identically.
 * The reference compiles with nvcc defaults (FMA contraction, __expf); its last-ulp
 * bits cannot be observed without CUDA, so integer decisions (sample counts,
 * cells, early-stop indices) are pinned against this convention.
 *
 * Sequential semantics: where the reference assigns output ranges with atomicAdd
 * in scheduling order (marching.cu:205, marching.cu:300), this restatement
 * processes rays / slots in index order, which is the order the HIP kernels
 * reproduce with prefix sums.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define SQRT3_F 1.7320508075688772f /* volrend.h:18 */
#define T_THRESHOLD 1e-4f           /* integrating.cu:12 */

/* ------------------------------------------------------------------------- */
/* deterministic exp                                                          */
/* ------------------------------------------------------------------------- */

/* exp(x) = 2^k * exp(r), k = rint(x / ln2), r = x - k ln2 with the two-constant
 * reduction of Cody & Waite (ln2 split into a 16-bit head and a tail), exp(r) by
 * its Taylor polynomial of degree 7 in Horner form (|r| <= ln2/2, truncation
 * error < 6e-9 relative).  2^k is applied as two exact power-of-two factors so
 * that subnormal results are rounded once. */
float ngp_expf(float x) {
    if (x != x) return x;
    if (x < -104.0f) return 0.0f;
    if (x > 88.72283935546875f) return INFINITY;
    float k = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(k, -0.693145751953125f, x);
    r = fmaf(k, -1.428606765330187045e-06f, r);
    float p = 1.98412701e-04f; /* 1/5040 */
    p = fmaf(p, r, 1.38888892e-03f); /* 1/720 */
    p = fmaf(p, r, 8.33333377e-03f); /* 1/120 */
    p = fmaf(p, r, 4.16666679e-02f); /* 1/24 */
    p = fmaf(p, r, 1.66666672e-01f); /* 1/6 */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int ki = (int)k;
    int k1 = ki / 2;
    int k2 = ki - k1;
    union { uint32_t u; float f; } s1, s2;
    s1.u = (uint32_t)(k1 + 127) << 23;
    s2.u = (uint32_t)(k2 + 127) << 23;
    return (p * s1.f) * s2.f;
}

/* ------------------------------------------------------------------------- */
/* Morton code (marching.cu:52-77)                                            */
/* ------------------------------------------------------------------------- */

/* Bit interleave by multiply-and-mask (the standard 10-bit spread; marching.cu:52-68).
 * For coordinates below 1024 it equals placing bit i of v on bit 3i. */
static uint32_t spread3(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
/* Inverse: keep every third bit and compact (marching.cu:70-77). */
static uint32_t gather3(uint32_t x) {
    x = x & 0x49249249u;
    x = (x | (x >> 2)) & 0xc30c30c3u;
    x = (x | (x >> 4)) & 0x0f00f00fu;
    x = (x | (x >> 8)) & 0xff0000ffu;
    x = (x | (x >> 16)) & 0x0000ffffu;
    return x;
}
uint32_t oracle_morton3d_one(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

void oracle_morton3d(uint32_t length, const uint32_t *xyzs, uint32_t *idcs) {
    for (uint32_t i = 0; i < length; ++i)
        idcs[i] = oracle_morton3d_one(xyzs[3 * i], xyzs[3 * i + 1], xyzs[3 * i + 2]);
}

void oracle_morton3d_invert(uint32_t length, const uint32_t *idcs, uint32_t *xyzs) {
    for (uint32_t i = 0; i < length; ++i) {
        xyzs[3 * i + 0] = gather3(idcs[i] >> 0);
        xyzs[3 * i + 1] = gather3(idcs[i] >> 1);
        xyzs[3 * i + 2] = gather3(idcs[i] >> 2);
    }
}

/* ------------------------------------------------------------------------- */
/* packbits (packbits.cu:9-34)                                                */
/* ------------------------------------------------------------------------- */

void oracle_packbits(uint32_t n_bytes, const float *thr, const float *density,
                     uint8_t *mask, uint8_t *bitfield) {
    for (uint32_t i = 0; i < n_bytes; ++i) {
        uint8_t byte = 0;
        for (int b = 0; b < 8; ++b) {
            int on = density[8 * i + b] > thr[8 * i + b];
            mask[8 * i + b] = (uint8_t)on;
            byte |= (uint8_t)(on << b);
        }
        bitfield[i] = byte;
    }
}

/* ------------------------------------------------------------------------- */
/* ray marching (marching.cu:14-397)                                          */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint32_t diag, K, G;
    float bound, portion, inv_G, ds_lo, ds_hi;
} march_ctx;

static march_ctx make_ctx(uint32_t diag, uint32_t K, uint32_t G, float bound, float portion) {
    march_ctx c;
    c.diag = diag; c.K = K; c.G = G; c.bound = bound; c.portion = portion;
    c.inv_G = 1.0f / (float)G;
    /* marching.cu:18-22: clamp(t*portion, 2*sqrt3*min(bound,1)/diag, 2*sqrt3*bound/G) */
    c.ds_lo = ((2.0f * SQRT3_F) * fminf(bound, 1.0f)) / (float)diag;
    c.ds_hi = ((2.0f * SQRT3_F) * bound) * c.inv_G;
    return c;
}

static float step_size(const march_ctx *c, float t) {
    return fminf(fmaxf(t * c->portion, c->ds_lo), c->ds_hi);
}

/* frexp exponent of |x| (marching.cu:25-39), computed from the bits */
static int frexp_exponent(float x) {
    int e;
    frexpf(x, &e);
    return e;
}

static uint32_t clamp_u(int v, int lo, int hi) { return (uint32_t)(v < lo ? lo : (v > hi ? hi : v)); }

/* Cell lookup of one position (marching.cu:79-98, 170-176).  Returns the
 * occupancy bit and fills the float grid coordinate and mip bound. */
static int lookup_cell(const march_ctx *c, const uint8_t *bitfield, const float pos[3], float ds,
                       float grid_f[3], float *mip_bound) {
    uint32_t cascade = 0;
    if (c->K > 1) {
        float m = fmaxf(fabsf(pos[0]), fmaxf(fabsf(pos[1]), fabsf(pos[2])));
        int e1 = frexp_exponent(m);
        int e2 = frexp_exponent(ds * (float)c->G);
        uint32_t c1 = clamp_u(e1, 0, (int)c->K - 1);
        uint32_t c2 = clamp_u(e2, 0, (int)c->K - 1);
        cascade = c1 > c2 ? c1 : c2;
    }
    *mip_bound = fminf(ldexpf(1.0f, (int)cascade), c->bound);
    uint32_t gi[3];
    for (int a = 0; a < 3; ++a) {
        grid_f[a] = ((pos[a] / *mip_bound + 1.0f) * 0.5f) * (float)c->G;
        /* clamp in float: equals clampi((int)floorf(.)) for every input, NaN -> 0 */
        gi[a] = (uint32_t)fminf(fmaxf(floorf(grid_f[a]), 0.0f), (float)(c->G - 1));
    }
    uint32_t idx = cascade * (c->G * c->G * c->G) + oracle_morton3d_one(gi[0], gi[1], gi[2]);
    return (bitfield[idx >> 3] >> (idx & 7u)) & 1u;
}

/* Next t after position `pos` in an empty cell: step by ds until past the cell's
 * exit along the ray (marching.cu:182-189).  A step that no longer changes t (t so
 * large that t + ds == t, where the reference loops forever) ends the march:
 * returns INFINITY. */
static float skip_empty(const march_ctx *c, float t, float ds, const float pos[3],
                        const float grid_f[3], float mip_bound, const float inv_d[3], const float d[3]) {
    float delta[3];
    for (int a = 0; a < 3; ++a) {
        float sgn = copysignf(1.0f, d[a]);
        float ng = floorf((grid_f[a] + 0.5f) + 0.5f * sgn);
        delta[a] = ((((ng * c->inv_G) - 0.5f) * 2.0f) * mip_bound - pos[a]) * inv_d[a];
    }
    float next_t = t + fmaxf(0.0f, fminf(delta[0], fminf(delta[1], delta[2])));
    float nt = t + ds;
    if (nt == t) return INFINITY;
    while (nt < next_t) {
        float s = nt + step_size(c, nt);
        if (s == nt) return INFINITY;
        nt = s;
    }
    return nt;
}

/* One training-march pass over a ray.  With `out_*` NULL only counts. */
static uint32_t march_one(const march_ctx *c, const float o[3], const float d[3], float t0, float t1,
                          float noise, const uint8_t *bitfield, uint32_t max_samples, uint32_t ray_index,
                          uint32_t *out_idcs, float *out_xyzs, float *out_dirs, float *out_dss, float *out_z) {
    float inv_d[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    float t = t0;
    t = t + step_size(c, t) * noise;
    uint32_t n = 0;
    float limit = (float)c->diag * c->bound;
    while ((float)n < limit && n < max_samples && t < t1) {
        float pos[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
        float ds = step_size(c, t);
        float gf[3], mb;
        int occ = lookup_cell(c, bitfield, pos, ds, gf, &mb);
        float nt;
        if (occ) {
            if (out_idcs) {
                out_idcs[n] = ray_index;
                for (int a = 0; a < 3; ++a) { out_xyzs[3 * n + a] = pos[a]; out_dirs[3 * n + a] = d[a]; }
                out_dss[n] = ds;
                out_z[n] = t;
            }
            ++n;
            nt = t + ds;
        } else {
            nt = skip_empty(c, t, ds, pos, gf, mb, inv_d, d);
        }
        t = nt;
    }
    return n;
}

/* march_rays (marching.cu:101-268 + wrapper marching/__init__.py:70-93).
 * Rays are processed in index order: ray i is valid iff the running sample count
 * before it is below total_samples and its own samples fit; a ray with no samples
 * is valid (unless t_end <= t_start).  Outputs are fully written. */
void oracle_march_rays(uint32_t n_rays, uint32_t total_samples, uint32_t diag, uint32_t K, uint32_t G,
                       float bound, float portion,
                       const float *rays_o, const float *rays_d, const float *t_starts, const float *t_ends,
                       const float *noises, const uint8_t *bitfield,
                       uint32_t *next_sample_write_location, uint32_t *number_of_exceeded_samples,
                       uint8_t *ray_is_valid, uint32_t *rays_n_samples, uint32_t *rays_sample_startidx,
                       uint32_t *idcs, float *xyzs, float *dirs, float *dss, float *z_vals) {
    march_ctx c = make_ctx(diag, K, G, bound, portion);
    memset(ray_is_valid, 0, n_rays);
    memset(rays_n_samples, 0, sizeof(uint32_t) * n_rays);
    memset(rays_sample_startidx, 0, sizeof(uint32_t) * n_rays);
    memset(idcs, 0, sizeof(uint32_t) * total_samples);
    memset(xyzs, 0, sizeof(float) * 3 * total_samples);
    memset(dirs, 0, sizeof(float) * 3 * total_samples);
    memset(dss, 0, sizeof(float) * total_samples);
    memset(z_vals, 0, sizeof(float) * total_samples);
    /* pass 1: counts (independent per ray) */
    uint32_t *counts = rays_n_samples;
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < (long)n_rays; ++i) {
        const float *o = rays_o + 3 * i, *d = rays_d + 3 * i;
        if (!(t_ends[i] > t_starts[i])) { counts[i] = 0; continue; }
        counts[i] = march_one(&c, o, d, t_starts[i], t_ends[i], noises[i], bitfield, 0xffffffffu, (uint32_t)i,
                              NULL, NULL, NULL, NULL, NULL);
    }
    /* sequential assignment */
    uint64_t written = 0, exceeded = 0;
    uint32_t counter = 0;
    int full = 0;
    for (uint32_t i = 0; i < n_rays; ++i) {
        uint32_t n = counts[i];
        counts[i] = 0;
        if (full || counter >= total_samples) { full = 1; continue; }
        if (!(t_ends[i] > t_starts[i])) continue;
        if (n == 0) { ray_is_valid[i] = 1; continue; }
        uint32_t start = counter;
        counter += n;
        if ((uint64_t)start + n > total_samples) { exceeded += n; full = 1; continue; }
        rays_sample_startidx[i] = start;
        counts[i] = n;
        ray_is_valid[i] = 1;
        written += n;
    }
    *next_sample_write_location = counter;
    *number_of_exceeded_samples = (uint32_t)exceeded;
    (void)written;
    /* pass 2: write samples of valid rays */
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < (long)n_rays; ++i) {
        uint32_t n = counts[i];
        if (n == 0) continue;
        uint32_t s = rays_sample_startidx[i];
        march_one(&c, rays_o + 3 * i, rays_d + 3 * i, t_starts[i], t_ends[i], noises[i], bitfield, n, (uint32_t)i,
                  idcs + s, xyzs + 3 * (size_t)s, dirs + 3 * (size_t)s, dss + s, z_vals + s);
    }
}

/* march_rays_inference (marching.cu:271-397).  Terminated slots take new ray
 * indices in slot order starting at next_ray_index_in. */
void oracle_march_rays_inference(uint32_t n_total_rays, uint32_t n_rays, uint32_t diag, uint32_t K, uint32_t G,
                                 uint32_t cap, float bound, float portion,
                                 const float *rays_o, const float *rays_d, const float *t_starts,
                                 const float *t_ends, const uint8_t *bitfield, const uint32_t *next_ray_index_in,
                                 const uint8_t *terminated, const uint32_t *indices_in,
                                 uint32_t *next_ray_index, uint32_t *indices_out, uint32_t *n_samples,
                                 float *t_starts_out, float *xyzs, float *dss, float *z_vals) {
    march_ctx c = make_ctx(diag, K, G, bound, portion);
    uint32_t next = next_ray_index_in[0];
    for (uint32_t i = 0; i < n_rays; ++i) indices_out[i] = terminated[i] ? next++ : indices_in[i];
    *next_ray_index = next;
    memset(n_samples, 0, sizeof(uint32_t) * n_rays);
    memset(t_starts_out, 0, sizeof(float) * n_rays);
    memset(xyzs, 0, sizeof(float) * 3 * (size_t)n_rays * cap);
    memset(dss, 0, sizeof(float) * (size_t)n_rays * cap);
    memset(z_vals, 0, sizeof(float) * (size_t)n_rays * cap);
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n_rays; ++i) {
        uint32_t r = indices_out[i];
        if (r >= n_total_rays) continue;
        const float *o = rays_o + 3 * (size_t)r, *d = rays_d + 3 * (size_t)r;
        float t0 = t_starts[r], t1 = t_ends[r];
        if (t1 < t0) continue;
        float *rx = xyzs + (size_t)i * cap * 3, *rds = dss + (size_t)i * cap, *rz = z_vals + (size_t)i * cap;
        float inv_d[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
        uint32_t steps = 0;
        float t = t0;
        while (steps < cap && t < t1) {
            float pos[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
            float ds = step_size(&c, t);
            float gf[3], mb;
            int occ = lookup_cell(&c, bitfield, pos, ds, gf, &mb);
            float nt;
            if (occ) {
                for (int a = 0; a < 3; ++a) rx[3 * steps + a] = pos[a];
                rds[steps] = ds;
                rz[steps] = t;
                ++steps;
                nt = t + ds;
            } else {
                nt = skip_empty(&c, t, ds, pos, gf, mb, inv_d, d);
            }
            t = nt;
        }
        if (t >= t1) {
            /* far-plane sample (marching.cu:367-394) */
            float pos[3] = {o[0] + t1 * d[0], o[1] + t1 * d[1], o[2] + t1 * d[2]};
            float ds = step_size(&c, t1);
            float gf[3], mb;
            if (lookup_cell(&c, bitfield, pos, ds, gf, &mb)) {
                if (steps > 0 && rds[steps - 1] + rz[steps - 1] >= t1) rds[steps - 1] = t1 - rz[steps - 1];
                if (steps < cap) {
                    for (int a = 0; a < 3; ++a) rx[3 * steps + a] = pos[a];
                    rds[steps] = ds;
                    rz[steps] = t1;
                    ++steps;
                } else {
                    t = t1;
                }
            }
        }
        n_samples[i] = steps;
        t_starts_out[i] = t;
    }
}

/* ------------------------------------------------------------------------- */
/* integration (integrating.cu:24-322)                                        */
/* ------------------------------------------------------------------------- */

/* samples of a ray inside the arrays (ranges past total_samples are cut there) */
static uint32_t clamp_count(uint32_t s0, uint32_t n, uint32_t total) {
    if (s0 >= total) return 0;
    return (uint64_t)s0 + n > total ? total - s0 : n;
}

void oracle_integrate_rays(uint32_t n_rays, uint32_t total_samples, const uint32_t *startidx,
                           const uint32_t *nsamp, const float *bgs, const float *dss, const float *z_vals,
                           const float *drgbs, uint32_t *measured_batch_size, float *final_rgbds,
                           float *final_opacities) {
    uint64_t measured = 0;
#pragma omp parallel for schedule(static) reduction(+ : measured)
    for (long i = 0; i < (long)n_rays; ++i) {
        uint32_t s0 = startidx[i], n = clamp_count(s0, nsamp[i], total_samples);
        float T = 1.0f, r = 0.0f, g = 0.0f, b = 0.0f, depth = 0.0f;
        uint32_t k = 0;
        for (; T > T_THRESHOLD && k < n; ++k) {
            size_t s = (size_t)s0 + k;
            float alpha = 1.0f - ngp_expf(-drgbs[4 * s] * dss[s]);
            float w = T * alpha;
            r = r + w * drgbs[4 * s + 1];
            g = g + w * drgbs[4 * s + 2];
            b = b + w * drgbs[4 * s + 3];
            depth = depth + w * z_vals[s];
            T = T * (1.0f - alpha);
        }
        float op = 1.0f - T;
        final_opacities[i] = op;
        if (T <= T_THRESHOLD) {
            float inv = 1.0f / op;
            final_rgbds[4 * i + 0] = r * inv;
            final_rgbds[4 * i + 1] = g * inv;
            final_rgbds[4 * i + 2] = b * inv;
            final_rgbds[4 * i + 3] = depth * inv;
        } else {
            final_rgbds[4 * i + 0] = r + T * bgs[3 * i + 0];
            final_rgbds[4 * i + 1] = g + T * bgs[3 * i + 1];
            final_rgbds[4 * i + 2] = b + T * bgs[3 * i + 2];
            final_rgbds[4 * i + 3] = depth;
        }
        measured += k;
    }
    *measured_batch_size = (uint32_t)measured;
}

void oracle_integrate_rays_backward(uint32_t n_rays, uint32_t total_samples, float near_distance,
                                    const uint32_t *startidx, const uint32_t *nsamp, const float *bgs,
                                    const float *dss, const float *z_vals, const float *drgbs,
                                    const float *final_rgbds, const float *final_opacities,
                                    const float *dL_dfinal_rgbds, float *dL_dbgs, float *dL_dz_vals,
                                    float *dL_ddrgbs) {
    memset(dL_dbgs, 0, sizeof(float) * 3 * n_rays);
    memset(dL_dz_vals, 0, sizeof(float) * total_samples);
    memset(dL_ddrgbs, 0, sizeof(float) * 4 * (size_t)total_samples);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n_rays; ++i) {
        uint32_t s0 = startidx[i], n = clamp_count(s0, nsamp[i], total_samples);
        const float *R = final_rgbds + 4 * i, *gr = dL_dfinal_rgbds + 4 * i, *bg = bgs + 3 * i;
        float op = final_opacities[i];
        int terminated = op >= 1.0f - T_THRESHOLD;
        float bgk = terminated ? 0.0f : 1.0f - op;
        float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f, cd = 0.0f;
        for (uint32_t k = 0; T > T_THRESHOLD && k < n; ++k) {
            size_t s = (size_t)s0 + k;
            float z = z_vals[s], ds = dss[s], sigma = drgbs[4 * s];
            float alpha = 1.0f - ngp_expf(-sigma * ds);
            float w = T * alpha;
            cr = cr + w * drgbs[4 * s + 1];
            cg = cg + w * drgbs[4 * s + 2];
            cb = cb + w * drgbs[4 * s + 3];
            cd = cd + w * z;
            T = T * (1.0f - alpha);
            dL_dz_vals[s] = w * gr[3];
            float a0 = gr[0] * ((T * drgbs[4 * s + 1] - (R[0] - cr)) - bg[0] * bgk);
            float a1 = gr[1] * ((T * drgbs[4 * s + 2] - (R[1] - cg)) - bg[1] * bgk);
            float a2 = gr[2] * ((T * drgbs[4 * s + 3] - (R[2] - cb)) - bg[2] * bgk);
            float a3 = gr[3] * (T * z - (R[3] - cd));
            float dsig = ds * (((a0 + a1) + a2) + a3);
            float reg = ((double)sigma > 4e-5 && z < near_distance) ? 1e-4f : 0.0f;
            float gs = fminf(z * z, 1.0f);
            dL_ddrgbs[4 * s + 0] = gs * dsig + reg;
            dL_ddrgbs[4 * s + 1] = w * gr[0];
            dL_ddrgbs[4 * s + 2] = w * gr[1];
            dL_ddrgbs[4 * s + 3] = w * gr[2];
        }
        if (T > T_THRESHOLD) {
            dL_dbgs[3 * i + 0] = T * gr[0];
            dL_dbgs[3 * i + 1] = T * gr[1];
            dL_dbgs[3 * i + 2] = T * gr[2];
        }
    }
}

void oracle_integrate_rays_inference(uint32_t n_total_rays, uint32_t n_rays, uint32_t cap, const float *rays_bg,
                                     const float *rays_rgbd, const float *rays_T, const uint32_t *n_samples,
                                     const uint32_t *indices, const float *dss, const float *z_vals,
                                     const float *drgbs, uint32_t *terminate_cnt, uint8_t *terminated,
                                     float *rgbd_out, float *T_out) {
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n_rays; ++i) {
        uint32_t n = n_samples[i] < cap ? n_samples[i] : cap, r = indices[i];
        terminated[i] = 0;
        for (int a = 0; a < 4; ++a) rgbd_out[4 * i + a] = 0.0f;
        T_out[i] = 0.0f;
        if (r >= n_total_rays) continue;
        const float *rds = dss + (size_t)i * cap, *rz = z_vals + (size_t)i * cap, *rc = drgbs + (size_t)i * cap * 4;
        float T = rays_T[r];
        float cr = rays_rgbd[4 * (size_t)r + 0], cg = rays_rgbd[4 * (size_t)r + 1];
        float cb = rays_rgbd[4 * (size_t)r + 2], cd = rays_rgbd[4 * (size_t)r + 3];
        for (uint32_t k = 0; T > T_THRESHOLD && k < n; ++k) {
            float alpha = 1.0f - ngp_expf(-rc[4 * k] * rds[k]);
            float w = T * alpha;
            cr = cr + w * rc[4 * k + 1];
            cg = cg + w * rc[4 * k + 2];
            cb = cb + w * rc[4 * k + 3];
            cd = cd + w * rz[k];
            T = T * (1.0f - alpha);
        }
        if (T <= T_THRESHOLD) {
            float inv = 1.0f / (1.0f - T);
            terminated[i] = 1;
            T_out[i] = 0.0f;
            rgbd_out[4 * i + 0] = cr * inv;
            rgbd_out[4 * i + 1] = cg * inv;
            rgbd_out[4 * i + 2] = cb * inv;
            rgbd_out[4 * i + 3] = cd * inv;
        } else {
            int term = n < cap;
            terminated[i] = (uint8_t)term;
            rgbd_out[4 * i + 3] = cd;
            T_out[i] = T;
            if (term) {
                rgbd_out[4 * i + 0] = cr + T * rays_bg[3 * (size_t)r + 0];
                rgbd_out[4 * i + 1] = cg + T * rays_bg[3 * (size_t)r + 1];
                rgbd_out[4 * i + 2] = cb + T * rays_bg[3 * (size_t)r + 2];
            } else {
                rgbd_out[4 * i + 0] = cr;
                rgbd_out[4 * i + 1] = cg;
                rgbd_out[4 * i + 2] = cb;
            }
        }
        cnt += terminated[i];
    }
    *terminate_cnt = cnt;
}

/* ------------------------------------------------------------------------- */
/* hash grid (models/encoders.py:82-233), f32                                 */
/* ------------------------------------------------------------------------- */

/* Vertex indices and weights of one point at one level.  Corner c = (cx,cy,cz)
 * with cz the fastest-varying bit (encoders.py:23-32). */
void oracle_hashgrid_level(const float pos[3], float bound, float scale, uint32_t res, int dense, uint32_t T,
                           uint32_t offset, uint32_t n_rows, uint32_t idx[8], float w[8]) {
    float fl[3], fr[3];
    for (int a = 0; a < 3; ++a) {
        float p = (pos[a] + bound) / (2.0f * bound);
        float ps = p * scale + 0.5f;
        fl[a] = floorf(ps);
        fr[a] = ps - fl[a];
    }
    for (int c = 0; c < 8; ++c) {
        uint32_t v[3];
        float wc = 1.0f;
        for (int a = 0; a < 3; ++a) {
            int o = (c >> (2 - a)) & 1;
            v[a] = (uint32_t)(fl[a] + (float)o);
            float wa = o ? fr[a] : 1.0f - fr[a];
            wc = a == 0 ? wa : wc * wa;
        }
        uint32_t h;
        if (dense) h = v[0] + v[1] * res + v[2] * (res * res);
        else h = v[0] ^ (v[1] * 2654435761u) ^ (v[2] * 805459861u);
        h = h % T + offset;
        if (h >= n_rows) h = n_rows - 1; /* gather indices are clamped (JAX semantics) */
        idx[c] = h;
        w[c] = wc;
    }
}

void oracle_hashgrid_fwd(uint32_t N, uint32_t L, uint32_t F, uint32_t T, uint32_t first_hash_level, uint32_t n_rows,
                         float bound, const float *scales, const uint32_t *res, const uint32_t *offsets,
                         const float *pos, const float *table, float *enc) {
#pragma omp parallel for schedule(static)
    for (long n = 0; n < (long)N; ++n) {
        for (uint32_t l = 0; l < L; ++l) {
            uint32_t idx[8];
            float w[8];
            oracle_hashgrid_level(pos + 3 * n, bound, scales[l], res[l], l < first_hash_level, T, offsets[l], n_rows,
                                  idx, w);
            for (uint32_t f = 0; f < F; ++f) {
                float acc = 0.0f;
                for (int c = 0; c < 8; ++c) acc = acc + table[(size_t)idx[c] * F + f] * w[c];
                enc[(size_t)n * L * F + l * F + f] = acc;
            }
        }
    }
}

/* Table gradient, accumulated in double. */
void oracle_hashgrid_bwd(uint32_t N, uint32_t L, uint32_t F, uint32_t T, uint32_t first_hash_level, uint32_t n_rows,
                         float bound, const float *scales, const uint32_t *res, const uint32_t *offsets,
                         const float *pos, const float *dL_denc, double *dL_dtable) {
    memset(dL_dtable, 0, sizeof(double) * (size_t)n_rows * F);
    for (uint32_t n = 0; n < N; ++n) {
        for (uint32_t l = 0; l < L; ++l) {
            uint32_t idx[8];
            float w[8];
            oracle_hashgrid_level(pos + 3 * n, bound, scales[l], res[l], l < first_hash_level, T, offsets[l], n_rows,
                                  idx, w);
            for (uint32_t f = 0; f < F; ++f) {
                double g = dL_denc[(size_t)n * L * F + l * F + f];
                for (int c = 0; c < 8; ++c) dL_dtable[(size_t)idx[c] * F + f] += (double)w[c] * g;
            }
        }
    }
}
