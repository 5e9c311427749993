// packbits_morton.hip — occupancy bitfield packing and Morton (z-order) codes.
//
// pack_density_into_bits: deps/volume-rendering-jax/lib/impl/packbits.cu:9-72
// morton3d / morton3d_invert: deps/volume-rendering-jax/lib/impl/marching.cu:52-77, 399-433, 606-665
#include "ngp_common.h"
#include "ngp_device.h"

namespace ngp {
namespace {

constexpr int kBlock = 256;

// One thread per output byte: 8 cells, bit k of byte j is cell 8j+k.  The 8
// densities and thresholds are read as two 16-byte vectors each and the 8 mask
// bytes are written as one 8-byte store.
__global__ void __launch_bounds__(kBlock) pack_bits_kernel(uint32_t n_bytes, const float *__restrict__ thr,
                                                           const float *__restrict__ density,
                                                           uint8_t *__restrict__ mask, uint8_t *__restrict__ bits,
                                                           bool vector_ok) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_bytes) return;
    float d[8], t[8];
    if (vector_ok) {
        const float4 *d4 = reinterpret_cast<const float4 *>(density) + 2 * (size_t)i;
        const float4 *t4 = reinterpret_cast<const float4 *>(thr) + 2 * (size_t)i;
        float4 a = d4[0], b = d4[1], c = t4[0], e = t4[1];
        d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w; d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
        t[0] = c.x; t[1] = c.y; t[2] = c.z; t[3] = c.w; t[4] = e.x; t[5] = e.y; t[6] = e.z; t[7] = e.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            d[k] = density[8 * (size_t)i + k];
            t[k] = thr[8 * (size_t)i + k];
        }
    }
    uint32_t byte = 0;
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t on = d[k] > t[k] ? 1u : 0u;
        byte |= on << k;
        m |= (uint64_t)on << (8 * k);
    }
    if (vector_ok) {
        reinterpret_cast<uint64_t *>(mask)[i] = m;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) mask[8 * (size_t)i + k] = (uint8_t)((m >> (8 * k)) & 1u);
    }
    bits[i] = (uint8_t)byte;
}

__global__ void __launch_bounds__(kBlock) morton3d_kernel(uint32_t length, const uint32_t *__restrict__ xyzs,
                                                          uint32_t *__restrict__ idcs) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= length) return;
    idcs[i] = morton3(xyzs[3 * (size_t)i], xyzs[3 * (size_t)i + 1], xyzs[3 * (size_t)i + 2]);
}

__global__ void __launch_bounds__(kBlock) morton3d_invert_kernel(uint32_t length, const uint32_t *__restrict__ idcs,
                                                                 uint32_t *__restrict__ xyzs) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= length) return;
    const uint32_t v = idcs[i];
    xyzs[3 * (size_t)i + 0] = morton_compact(v);
    xyzs[3 * (size_t)i + 1] = morton_compact(v >> 1);
    xyzs[3 * (size_t)i + 2] = morton_compact(v >> 2);
}

inline bool aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

}  // namespace
}  // namespace ngp

using namespace ngp;

extern "C" void pack_density_into_bits(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    PackbitsDescriptor desc;
    if (!read_descriptor(opaque, opaque_len, &desc, "pack_density_into_bits")) return;
    if (desc.n_bytes == 0) return;
    if (!check_buffers(buffers, 4, "pack_density_into_bits")) return;
    const float *thr = static_cast<const float *>(buffers[0]);
    const float *density = static_cast<const float *>(buffers[1]);
    uint8_t *mask = static_cast<uint8_t *>(buffers[2]);
    uint8_t *bits = static_cast<uint8_t *>(buffers[3]);
    bool vec = aligned(thr, 16) && aligned(density, 16) && aligned(mask, 8);
    hipLaunchKernelGGL(pack_bits_kernel, dim3(div_up(desc.n_bytes, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, desc.n_bytes, thr, density, mask, bits, vec);
    check_launch("pack_density_into_bits");
}

extern "C" void morton3d(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    Morton3DDescriptor desc;
    if (!read_descriptor(opaque, opaque_len, &desc, "morton3d")) return;
    if (desc.length == 0) return;
    if (!check_buffers(buffers, 2, "morton3d")) return;
    hipLaunchKernelGGL(morton3d_kernel, dim3(div_up(desc.length, kBlock)), dim3(kBlock), 0, (hipStream_t)stream,
                       desc.length, static_cast<const uint32_t *>(buffers[0]), static_cast<uint32_t *>(buffers[1]));
    check_launch("morton3d");
}

extern "C" void morton3d_invert(ngp_stream_t stream, void **buffers, const char *opaque, size_t opaque_len) {
    Morton3DDescriptor desc;
    if (!read_descriptor(opaque, opaque_len, &desc, "morton3d_invert")) return;
    if (desc.length == 0) return;
    if (!check_buffers(buffers, 2, "morton3d_invert")) return;
    hipLaunchKernelGGL(morton3d_invert_kernel, dim3(div_up(desc.length, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, desc.length, static_cast<const uint32_t *>(buffers[0]),
                       static_cast<uint32_t *>(buffers[1]));
    check_launch("morton3d_invert");
}
