"""ctypes binding of libngp_hip.so (the C ABI declared in include/ngp_hip.h).

Every kernel target has the legacy XLA custom-call signature
``void target(stream, void **buffers, const char *opaque, size_t opaque_len)``;
this module packs descriptors (byte-identical to the reference's
deps/volume-rendering-jax/lib/impl/volrend.h:23-118 structs), builds the buffer
array from torch tensors, calls the target on the current HIP stream and turns a
recorded error into a Python exception.

There is no fallback: if the library is missing, importing a kernel wrapper
raises.  Build it with ``python -m jaxngp_amd.build``.
"""
from __future__ import annotations

import ctypes
import struct
from pathlib import Path
from typing import Sequence

import torch

LIB_PATH = Path(__file__).resolve().parent / "libngp_hip.so"

TARGETS = (
    "march_rays",
    "march_rays_inference",
    "morton3d",
    "morton3d_invert",
    "pack_density_into_bits",
    "integrate_rays",
    "integrate_rays_backward",
    "integrate_rays_inference",
    "hashgrid_encode",
    "hashgrid_encode_backward",
    "tcnn_hashgrid_encode",
    "tcnn_hashgrid_encode_backward",
    "sky_sphere_bg",
    "sky_sphere_bg_backward",
    "nerf_mlp_forward",
    "nerf_mlp_backward",
    "nerf_mlp_backward_no_alloc",
    "train_make_rays",
    "train_gt_blend",
    "train_bg_grad",
    "train_loss_grad",
    "adam_update",
    "loss_scale_check",
    "loss_scale_update",
    "nerf_field_chunks",
    "occupancy_grid_update",
    "pool_scatter",
    "proposal_points",
    "proposal_mlp_forward",
    "proposal_mlp_backward",
    "proposal_loss",
    "proposal_sample",
    "proposal_render_samples",
    "proposal_train_samples",
    "hashgrid_encode_backward_adam",
)

HASHGRID_MAX_LEVELS = 32
DTYPE_F32 = 0
DTYPE_F16 = 1
MLP_N_PARAMS = 9408


class NGPError(RuntimeError):
    """An error recorded by libngp_hip.so (bad descriptor, bad buffer, launch failure)."""


_lib = None


def load() -> ctypes.CDLL:
    """Loads libngp_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP kernels with `python -m jaxngp_amd.build` "
            "(there is no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.ngp_get_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.ngp_get_last_error.restype = ctypes.c_int
    lib.ngp_version.argtypes = []
    lib.ngp_version.restype = ctypes.c_char_p
    lib.hashgrid_encode_backward_workspace_size.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.hashgrid_encode_backward_workspace_size.restype = ctypes.c_size_t
    lib.hashgrid_encode_backward_adam_supported.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.hashgrid_encode_backward_adam_supported.restype = ctypes.c_int
    lib.sky_sphere_bg_workspace_size.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.sky_sphere_bg_workspace_size.restype = ctypes.c_size_t
    lib.sky_sphere_bg_n_params.argtypes = [ctypes.c_uint32]
    lib.sky_sphere_bg_n_params.restype = ctypes.c_uint32
    _lib = lib
    return lib


_SIG = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_size_t]
_targets: dict = {}


def target(name: str):
    """The C function of kernel target `name`, with its signature set."""
    fn = _targets.get(name)
    if fn is None:
        fn = getattr(load(), name)
        fn.argtypes = _SIG
        fn.restype = None
        _targets[name] = fn
    return fn


def last_error() -> tuple[int, str]:
    buf = ctypes.create_string_buffer(512)
    code = load().ngp_get_last_error(buf, len(buf))
    return code, buf.value.decode()


def _check() -> None:
    code, msg = last_error()
    if code != 0:
        raise NGPError(msg)


def current_stream(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def call(name: str, tensors: Sequence[torch.Tensor], opaque: bytes, stream: int | None = None) -> None:
    """Calls target `name` with device tensors (operands then results)."""
    for i, t in enumerate(tensors):
        if not t.is_cuda:
            raise ValueError(f"{name}: buffer {i} is not a device tensor")
        if not t.is_contiguous():
            raise ValueError(f"{name}: buffer {i} is not contiguous")
    ptrs = (ctypes.c_void_p * max(1, len(tensors)))(*[t.data_ptr() for t in tensors])
    if stream is None:
        stream = current_stream(tensors[0].device if tensors else None)
    target(name)(ctypes.c_void_p(stream), ptrs, opaque, len(opaque))
    _check()


class Prepared:
    """A target call with its buffers bound once: the pointer array and the stream
    are built at construction, so a call costs one ctypes call.  The tensors are
    kept referenced; `set_buffer` swaps one.  Errors are left for `check()`."""

    __slots__ = ("name", "fn", "ptrs", "opaque", "stream", "keep")

    def __init__(self, name: str, tensors: Sequence[torch.Tensor], opaque: bytes, stream: int | None = None):
        for i, t in enumerate(tensors):
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name}: buffer {i} is not a contiguous device tensor")
        self.name = name
        self.fn = target(name)
        self.keep = list(tensors)
        self.ptrs = (ctypes.c_void_p * max(1, len(tensors)))(*[t.data_ptr() for t in tensors])
        self.opaque = opaque
        if stream is None:
            stream = current_stream(tensors[0].device if tensors else None)
        self.stream = ctypes.c_void_p(stream)

    def set_buffer(self, i: int, t: torch.Tensor) -> None:
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{self.name}: buffer {i} is not a contiguous device tensor")
        self.keep[i] = t
        self.ptrs[i] = t.data_ptr()

    def __call__(self, opaque: bytes | None = None, stream: int | None = None) -> None:
        """Launches with the bound descriptor (or `opaque`) on the bound stream (or
        `stream`, a raw stream handle)."""
        o = self.opaque if opaque is None else opaque
        self.fn(self.stream if stream is None else ctypes.c_void_p(stream), self.ptrs, o, len(o))


def check() -> None:
    """Raises NGPError if a target recorded an error on this thread."""
    _check()


def hashgrid_backward_workspace_bytes(opaque: bytes) -> int:
    """Workspace bytes of hashgrid_encode_backward for descriptor `opaque`."""
    n = load().hashgrid_encode_backward_workspace_size(opaque, len(opaque))
    if n == 0:
        _check()
        raise NGPError("hashgrid_encode_backward_workspace_size: invalid descriptor")
    return n


def hashgrid_backward_adam_supported(opaque: bytes) -> bool:
    """Whether hashgrid_encode_backward_adam takes the HashGridDescriptor `opaque`
    (F <= 2, f32 table, levels within the one-pass table gradient)."""
    ok = bool(load().hashgrid_encode_backward_adam_supported(opaque, len(opaque)))
    _check()
    return ok


def hashgrid_adam_descriptor(grid: bytes, adam: bytes) -> bytes:
    """HashGridAdamDescriptor: a HashGridDescriptor, then an NgpAdamDescriptor (the
    first is 424 bytes, a multiple of the second's 8-byte alignment)."""
    assert len(grid) % 8 == 0
    return grid + adam


def sky_sphere_bg_workspace_bytes(opaque: bytes) -> int:
    """Workspace bytes of sky_sphere_bg / sky_sphere_bg_backward for `opaque`."""
    n = load().sky_sphere_bg_workspace_size(opaque, len(opaque))
    if n == 0:
        _check()
        raise NGPError("sky_sphere_bg_workspace_size: invalid descriptor")
    return n


def call_raw(name: str, ptrs: Sequence[int], opaque: bytes, stream: int = 0) -> None:
    """Calls target `name` with raw pointers (host-side error tests use NULLs)."""
    arr = (ctypes.c_void_p * max(1, len(ptrs)))(*ptrs)
    target(name)(ctypes.c_void_p(stream), arr, opaque, len(opaque))
    _check()


# ---------------------------------------------------------------------------
# descriptors (little-endian, natural alignment, no padding: all 4-byte fields)
# ---------------------------------------------------------------------------

def marching_descriptor(n_rays, total_samples, diagonal_n_steps, K, G, bound, stepsize_portion) -> bytes:
    return struct.pack("<5I2f", n_rays, total_samples, diagonal_n_steps, K, G, bound, stepsize_portion)


def marching_inference_descriptor(n_total_rays, n_rays, diagonal_n_steps, K, G, march_steps_cap, bound,
                                  stepsize_portion) -> bytes:
    return struct.pack("<6I2f", n_total_rays, n_rays, diagonal_n_steps, K, G, march_steps_cap, bound,
                       stepsize_portion)


def integrating_descriptor(n_rays, total_samples) -> bytes:
    return struct.pack("<2I", n_rays, total_samples)


def integrating_backward_descriptor(n_rays, total_samples, near_distance) -> bytes:
    return struct.pack("<2If", n_rays, total_samples, near_distance)


def integrating_inference_descriptor(n_total_rays, n_rays, march_steps_cap) -> bytes:
    return struct.pack("<3I", n_total_rays, n_rays, march_steps_cap)


def morton3d_descriptor(length) -> bytes:
    return struct.pack("<I", length)


def packbits_descriptor(n_bytes) -> bytes:
    return struct.pack("<I", n_bytes)


ENC_POINT_MAJOR = 0
ENC_LEVEL_MAJOR = 1


def hashgrid_descriptor(n_points, L, F, T, first_hash_level, table_dtype, n_rows, bound, scales, resolutions,
                        offsets, enc_layout=ENC_POINT_MAJOR) -> bytes:
    M = HASHGRID_MAX_LEVELS
    if not (1 <= L <= M):
        raise ValueError(f"L={L} must be in [1, {M}]")
    sc = list(scales) + [0.0] * (M - L)
    rs = list(resolutions) + [0] * (M - L)
    of = list(offsets) + [offsets[-1]] * (M + 1 - len(offsets))
    return struct.pack(f"<7IfI{M}f{M}I{M + 1}I", n_points, L, F, T, first_hash_level, table_dtype, n_rows, bound,
                       enc_layout, *sc, *rs, *of)


def tcnn_hashgrid_descriptor(n_coords, L, F, N_min, per_level_scale) -> bytes:
    """jaxtcnn::HashGridDescriptor (deps/jax-tcnn/lib/impl/tcnnutils.h:11-29)."""
    return struct.pack("<4If", n_coords, L, F, N_min, per_level_scale)


def train_bg_descriptor(n_rays) -> bytes:
    return struct.pack("<I", n_rays)


SKY_MAX_EXTRA_DIMS = 16


def sky_sphere_bg_n_params(n_extra_dims: int) -> int:
    """(20 + E) * 32 + 32 * 32 + 32 * 3 flat f32 parameters of the sky-sphere MLP."""
    return (20 + int(n_extra_dims)) * 32 + 32 * 32 + 32 * 3


def sky_sphere_bg_descriptor(n_rays, n_extra_dims, appearance_per_ray, radius) -> bytes:
    """SkySphereBgDescriptor of include/ngp_hip.h."""
    return struct.pack("<3If", n_rays, n_extra_dims, int(appearance_per_ray), radius)


def nerf_mlp_descriptor(n_samples, n_workspace_blocks=0, enc_layout=ENC_POINT_MAJOR, param_dtype=DTYPE_F32) -> bytes:
    return struct.pack("<4I", n_samples, n_workspace_blocks, enc_layout, param_dtype)


TRAIN_PERTURB = 1
TRAIN_RANDOM_BG = 2


def train_rays_descriptor(n_rays, n_views, width, height, fx, fy, cx, cy, bound, seed, step, flags, bg) -> bytes:
    return struct.pack("<4I5f4I3f", n_rays, n_views, width, height, fx, fy, cx, cy, bound, seed & 0xFFFFFFFF,
                       (seed >> 32) & 0xFFFFFFFF, step & 0xFFFFFFFF, flags, *bg)


LOSS_SCALED = 1


def loss_descriptor(n_rays, delta, flags=0) -> bytes:
    return struct.pack("<IfI", n_rays, delta, flags)


def adam_descriptor(n, lr, b1, b2, eps, eps_root, b1_pow_t, b2_pow_t, weight_decay, grad_scale=1.0,
                    mirror_dtype=DTYPE_F32, flags=0) -> bytes:
    return struct.pack("<Q9f3I", n, lr, b1, b2, eps, eps_root, b1_pow_t, b2_pow_t, weight_decay, grad_scale,
                       mirror_dtype, flags, 0)


PROPOSAL_BINS = 64
PROPOSAL_N_PARAMS = 176


def proposal_rays_descriptor(n_rays) -> bytes:
    return struct.pack("<I", n_rays)


def proposal_mlp_descriptor(n_points, n_workspace_blocks=0, enc_layout=ENC_POINT_MAJOR, param_dtype=DTYPE_F32) -> bytes:
    return struct.pack("<4I", n_points, n_workspace_blocks, enc_layout, param_dtype)


PROPOSAL_QUANTILES = 0
PROPOSAL_SUPPORT = 1


def proposal_sample_descriptor(n_rays, n_per_ray=48, mass_threshold=0.98, mode=PROPOSAL_QUANTILES, eps=1e-3,
                               diagonal_n_steps=1024, K=1, G=128, bound=1.0, stepsize_portion=0.0,
                               empty_mass=0.0) -> bytes:
    return struct.pack("<2IfIf3I3f", n_rays, n_per_ray, mass_threshold, mode, eps, diagonal_n_steps, K, G, bound,
                       stepsize_portion, empty_mass)


def proposal_render_descriptor(grid_desc: bytes, sample_desc: bytes, param_dtype=DTYPE_F32) -> bytes:
    """ProposalRenderDescriptor: HashGridDescriptor, ProposalSampleDescriptor, u32."""
    return grid_desc + sample_desc + struct.pack("<I", param_dtype)


def proposal_train_descriptor(render_desc: bytes, total_samples, n_uniform=0) -> bytes:
    """ProposalTrainDescriptor: ProposalRenderDescriptor, u32 total_samples, u32 n_uniform."""
    return render_desc + struct.pack("<2I", total_samples, n_uniform)


def proposal_loss_descriptor(n_rays, eps=1e-7) -> bytes:
    return struct.pack("<If", n_rays, eps)


def pool_scatter_descriptor(n, n_total, width) -> bytes:
    return struct.pack("<3I", n, n_total, width)


def field_chunks_descriptor(grid_desc: bytes, n_slots, cap, n_total_rays, param_dtype=DTYPE_F32) -> bytes:
    """NerfFieldChunksDescriptor: a HashGridDescriptor followed by four u32."""
    return grid_desc + struct.pack("<4I", n_slots, cap, n_total_rays, param_dtype)


def occupancy_update_descriptor(grid_desc: bytes, G, K, cascade, n_alive, update_all, param_dtype, seed, step) -> bytes:
    """OccupancyGridUpdateDescriptor: a HashGridDescriptor followed by nine u32."""
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return grid_desc + struct.pack("<9I", G, K, cascade, n_alive, int(bool(update_all)), param_dtype,
                                   seed & 0xFFFFFFFF, seed >> 32, int(step) & 0xFFFFFFFF)


def loss_scale_check_descriptor(n) -> bytes:
    return struct.pack("<Q", n)


def loss_scale_update_descriptor(growth_interval=2000, min_scale=1.0, max_scale=2.0 ** 24) -> bytes:
    return struct.pack("<I2f", growth_interval, min_scale, max_scale)


def loss_scale_state(device, scale=2.0 ** 15) -> torch.Tensor:
    """NgpLossScaleState {scale f32, overflow u32, good_steps u32, reserved} as u8[16]."""
    return torch.tensor(list(struct.pack("<f3I", scale, 0, 0, 0)), dtype=torch.uint8, device=device)
