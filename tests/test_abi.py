"""C-ABI checks that need no GPU: the library loads, exports every entry point the
header declares, descriptor packing matches the C structs, and host-side failures
are reported through ngp_get_last_error instead of crashing (the reference throws
C++ exceptions across the XLA boundary, volrend.h:9-16, serde.h:35-41)."""
import re
import struct
import subprocess
from pathlib import Path

import pytest

from jaxngp_amd import _lib

REPO = Path(__file__).resolve().parents[1]
HEADER = REPO / "include" / "ngp_hip.h"


def _declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:void|int|size_t|uint32_t|const char \*)\s*\*?\s*([a-z_0-9]+)\s*\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_functions()
    assert "march_rays" in declared and "ngp_get_last_error" in declared
    assert "hashgrid_encode_backward_workspace_size" in declared
    for name in declared:
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in nm.splitlines() if line.strip()}
    assert set(declared) <= exported
    assert set(_lib.TARGETS) <= set(declared)


def test_version():
    assert _lib.load().ngp_version().decode().startswith("ngp-hip")


def test_descriptor_sizes_match_c_structs():
    # sizes of the structs in include/ngp_hip.h (reference volrend.h:23-118 for the first seven)
    assert len(_lib.integrating_descriptor(1, 2)) == 8
    assert len(_lib.integrating_backward_descriptor(1, 2, 0.3)) == 12
    assert len(_lib.integrating_inference_descriptor(1, 2, 3)) == 12
    assert len(_lib.marching_descriptor(1, 2, 3, 4, 5, 1.0, 0.0)) == 28
    assert len(_lib.marching_inference_descriptor(1, 2, 3, 4, 5, 6, 1.0, 0.0)) == 32
    assert len(_lib.morton3d_descriptor(5)) == 4
    assert len(_lib.packbits_descriptor(5)) == 4
    assert len(_lib.hashgrid_descriptor(1, 2, 2, 8, 1, 0, 16, 1.0, [1.0, 2.0], [2, 3], [0, 8, 16])) == 9 * 4 + 32 * 4 * 2 + 33 * 4
    assert len(_lib.nerf_mlp_descriptor(5, 0)) == 16
    # jaxtcnn::HashGridDescriptor (deps/jax-tcnn/lib/impl/tcnnutils.h:11-29): 4 u32 + f32
    assert len(_lib.tcnn_hashgrid_descriptor(7, 16, 2, 16, 1.38)) == 20
    assert len(_lib.sky_sphere_bg_descriptor(7, 0, 1, 4.0)) == 16
    assert len(_lib.train_rays_descriptor(1, 2, 3, 4, 1.0, 1.0, 1.0, 1.0, 1.0, 5, 6, 3, (1.0, 1.0, 1.0))) == 64
    assert len(_lib.loss_descriptor(4, 0.1)) == 12
    assert len(_lib.adam_descriptor(10, 1e-2, 0.9, 0.99, 1e-15, 1e-15, 0.9, 0.99, 0.0)) == 56
    assert len(_lib.loss_scale_check_descriptor(10)) == 8
    assert len(_lib.loss_scale_update_descriptor()) == 12
    assert len(_lib.pool_scatter_descriptor(1, 2, 3)) == 12
    assert len(_lib.proposal_rays_descriptor(1)) == 4
    assert len(_lib.proposal_mlp_descriptor(1, 2)) == 16
    assert len(_lib.proposal_loss_descriptor(1)) == 8
    assert len(_lib.proposal_sample_descriptor(1)) == 44
    hg5 = _lib.hashgrid_descriptor(1, 2, 2, 8, 1, 0, 16, 1.0, [1.0, 2.0], [2, 3], [0, 8, 16])
    assert len(_lib.proposal_render_descriptor(hg5, _lib.proposal_sample_descriptor(1))) == len(hg5) + 48
    hg = _lib.hashgrid_descriptor(1, 2, 2, 8, 1, 0, 16, 1.0, [1.0, 2.0], [2, 3], [0, 8, 16])
    assert len(_lib.field_chunks_descriptor(hg, 1, 8, 3)) == len(hg) + 16
    # field order: n_rays first, bound at byte 20
    d = _lib.marching_descriptor(7, 9, 1024, 1, 128, 1.5, 0.25)
    assert struct.unpack_from("<I", d, 0)[0] == 7 and struct.unpack_from("<f", d, 20)[0] == 1.5


def test_compile_header_as_c():
    """The header is plain C (no HIP types) and its struct sizes are as documented."""
    src = r'''
#include "ngp_hip.h"
_Static_assert(sizeof(MarchingDescriptor) == 28, "");
_Static_assert(sizeof(MarchingInferenceDescriptor) == 32, "");
_Static_assert(sizeof(IntegratingDescriptor) == 8, "");
_Static_assert(sizeof(IntegratingBackwardDescriptor) == 12, "");
_Static_assert(sizeof(IntegratingInferenceDescriptor) == 12, "");
_Static_assert(sizeof(PackbitsDescriptor) == 4, "");
_Static_assert(sizeof(Morton3DDescriptor) == 4, "");
_Static_assert(sizeof(HashGridDescriptor) == 9 * 4 + 32 * 8 + 33 * 4, "");
_Static_assert(sizeof(NerfMlpDescriptor) == 16, "");
_Static_assert(sizeof(NgpTrainRaysDescriptor) == 64, "");
_Static_assert(sizeof(NgpLossDescriptor) == 12, "");
_Static_assert(sizeof(NgpAdamDescriptor) == 56, "");
_Static_assert(sizeof(NgpLossScaleState) == 16, "");
_Static_assert(sizeof(NgpLossScaleCheckDescriptor) == 8, "");
_Static_assert(sizeof(NgpLossScaleUpdateDescriptor) == 12, "");
_Static_assert(sizeof(NgpPoolScatterDescriptor) == 12, "");
_Static_assert(sizeof(ProposalRaysDescriptor) == 4, "");
_Static_assert(sizeof(ProposalMlpDescriptor) == 16, "");
_Static_assert(sizeof(ProposalLossDescriptor) == 8, "");
_Static_assert(sizeof(ProposalSampleDescriptor) == 44, "");
_Static_assert(sizeof(ProposalRenderDescriptor) == sizeof(HashGridDescriptor) + 48, "");
_Static_assert(sizeof(NerfFieldChunksDescriptor) == sizeof(HashGridDescriptor) + 16, "");
_Static_assert(sizeof(TcnnHashGridDescriptor) == 20, "");
_Static_assert(sizeof(HashGridAdamDescriptor) == sizeof(HashGridDescriptor) + 56, "");
_Static_assert(sizeof(SkySphereBgDescriptor) == 16, "");
int main(void) { return 0; }
'''
    out = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", f"-I{REPO / 'include'}", "-x", "c", "-", "-o",
                          "/dev/null"], input=src, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr


@pytest.mark.parametrize("target", _lib.TARGETS)
def test_wrong_descriptor_size_is_an_error(target):
    with pytest.raises(_lib.NGPError, match="opaque"):
        _lib.call_raw(target, [0] * 16, b"\x00\x00\x00")


def test_null_buffer_is_an_error():
    with pytest.raises(_lib.NGPError, match="NULL"):
        _lib.call_raw("march_rays", [0] * 16, _lib.marching_descriptor(4, 16, 1024, 1, 128, 1.0, 0.0))
    with pytest.raises(_lib.NGPError, match="NULL"):
        _lib.call_raw("integrate_rays_backward", [0] * 12, _lib.integrating_backward_descriptor(4, 16, 0.3))


def test_bad_statics_are_errors():
    with pytest.raises(_lib.NGPError, match="statics"):
        _lib.call_raw("march_rays", [0] * 16, _lib.marching_descriptor(4, 16, 1024, 0, 128, 1.0, 0.0))
    with pytest.raises(_lib.NGPError, match="statics"):
        _lib.call_raw("march_rays", [0] * 16, _lib.marching_descriptor(4, 16, 1024, 1, 128, -1.0, 0.0))
    with pytest.raises(_lib.NGPError, match="total_samples"):
        _lib.call_raw("march_rays", [0] * 16, _lib.marching_descriptor(4, 0, 1024, 1, 128, 1.0, 0.0))
    with pytest.raises(_lib.NGPError, match="march_steps_cap"):
        _lib.call_raw("march_rays_inference", [0] * 15,
                      _lib.marching_inference_descriptor(10, 4, 1024, 1, 128, 0, 1.0, 0.0))
    with pytest.raises(_lib.NGPError, match="F="):
        _lib.call_raw("hashgrid_encode", [0] * 3, _lib.hashgrid_descriptor(4, 1, 3, 8, 1, 0, 8, 1.0, [1.0], [2], [0, 8]))
    with pytest.raises(_lib.NGPError, match="level table"):
        _lib.call_raw("hashgrid_encode", [0] * 3, _lib.hashgrid_descriptor(4, 1, 2, 8, 1, 0, 9, 1.0, [1.0], [2], [0, 8]))
    # the error is cleared once read
    assert _lib.last_error() == (0, "")


def test_hashgrid_backward_workspace_size():
    """Host-only query: the workspace holds entries for up to 8 rows per point and
    level (F=2: 12 bytes with an f32 table, 8 bytes with an f16 table), and an
    invalid descriptor is an error, not 0 bytes."""
    from jaxngp_amd.encoders import level_table
    lv = level_table(16, 2 ** 19, 16, 2048)
    desc = lambda n, dt=0: _lib.hashgrid_descriptor(n, 16, 2, 2 ** 19, lv.first_hash_level, dt, lv.n_rows, 1.0,
                                                    list(lv.scales), lv.resolutions, lv.offsets)
    for dt, entry in ((0, 12), (1, 8)):
        small = _lib.hashgrid_backward_workspace_bytes(desc(1000, dt))
        big = _lib.hashgrid_backward_workspace_bytes(desc(2 ** 18, dt))
        assert big - small >= (2 ** 18 - 1000) * 16 * 8 * entry
        assert big - small < (2 ** 18 - 1000) * 16 * 8 * (entry + 4)
        assert big < 2 ** 30
    with pytest.raises(_lib.NGPError, match="F="):
        _lib.hashgrid_backward_workspace_bytes(_lib.hashgrid_descriptor(4, 1, 3, 8, 1, 0, 8, 1.0, [1.0], [2], [0, 8]))
    assert _lib.last_error() == (0, "")


def test_sky_sphere_host_checks():
    """Parameter count, workspace size and descriptor errors need no GPU."""
    lib = _lib.load()
    for E in (0, 3, 16):
        assert lib.sky_sphere_bg_n_params(E) == _lib.sky_sphere_bg_n_params(E) == (20 + E) * 32 + 1120
    small = _lib.sky_sphere_bg_workspace_bytes(_lib.sky_sphere_bg_descriptor(10, 0, 1, 4.0))
    big = _lib.sky_sphere_bg_workspace_bytes(_lib.sky_sphere_bg_descriptor(1 << 20, 16, 1, 4.0))
    assert small % 256 == 0 and big % 256 == 0 and small < big <= 2304 + 512 * 2272 * 4 + 255
    for bad in (_lib.sky_sphere_bg_descriptor(10, 17, 1, 4.0), _lib.sky_sphere_bg_descriptor(10, 0, 2, 4.0),
                _lib.sky_sphere_bg_descriptor(10, 0, 1, 0.0), _lib.sky_sphere_bg_descriptor(10, 0, 1, float("nan"))):
        with pytest.raises(_lib.NGPError, match="invalid descriptor"):
            _lib.sky_sphere_bg_workspace_bytes(bad)
    with pytest.raises(_lib.NGPError, match="per-ray appearance"):
        _lib.call_raw("sky_sphere_bg_backward", [0] * 8, _lib.sky_sphere_bg_descriptor(10, 4, 0, 4.0))
    with pytest.raises(_lib.NGPError, match="NULL"):
        _lib.call_raw("sky_sphere_bg", [0] * 6, _lib.sky_sphere_bg_descriptor(10, 0, 1, 4.0))
    assert _lib.last_error() == (0, "")


def test_no_alloc_mlp_backward_host_checks():
    """The no-allocation MLP backward refuses more workgroup rows than its fixed
    scratch holds, and NULL buffers, before touching a device."""
    with pytest.raises(_lib.NGPError, match="exceeds the fixed scratch"):
        _lib.call_raw("nerf_mlp_backward_no_alloc", [16] * 6, _lib.nerf_mlp_descriptor(0, 513))
    with pytest.raises(_lib.NGPError, match="NULL"):
        _lib.call_raw("nerf_mlp_backward_no_alloc", [0] * 6, _lib.nerf_mlp_descriptor(64, 8))
    with pytest.raises(_lib.NGPError, match="n_workspace_blocks must be positive"):
        _lib.call_raw("nerf_mlp_backward_no_alloc", [16] * 6, _lib.nerf_mlp_descriptor(64, 0))
    assert _lib.last_error() == (0, "")


def test_error_is_seen_from_another_thread():
    """A host framework may run targets on its own executor threads (XLA does) while
    another thread polls ngp_get_last_error: the error state is process-wide."""
    import threading
    fn = _lib.target("march_rays")
    arr = (_lib.ctypes.c_void_p * 16)(*([None] * 16))
    bad = _lib.marching_descriptor(4, 0, 1024, 1, 128, 1.0, 0.0)  # total_samples = 0
    t = threading.Thread(target=lambda: fn(_lib.ctypes.c_void_p(0), arr, bad, len(bad)))
    t.start()
    t.join()
    code, msg = _lib.last_error()
    assert code != 0 and "total_samples" in msg
    assert _lib.last_error() == (0, "")
    # the first error recorded since the last poll is the one reported
    t1 = threading.Thread(target=lambda: fn(_lib.ctypes.c_void_p(0), arr, bad, len(bad)))
    t1.start()
    t1.join()
    _lib.target("integrate_rays")(_lib.ctypes.c_void_p(0), arr, b"\x00", 1)
    code, msg = _lib.last_error()
    assert "total_samples" in msg
    assert _lib.last_error() == (0, "")


def test_fused_table_update_host_checks():
    """hashgrid_encode_backward_adam: which descriptors it takes (F <= 2, f32 table,
    the one-pass table gradient), and that the others are errors before any launch."""
    from jaxngp_amd.encoders import level_table
    lv = level_table(16, 2 ** 19, 16, 2048)
    desc = lambda F=2, dt=0, T=2 ** 19, n=1000: _lib.hashgrid_descriptor(
        n, 16, F, T, lv.first_hash_level, dt, lv.n_rows, 1.0, list(lv.scales), lv.resolutions, lv.offsets)
    assert _lib.hashgrid_backward_adam_supported(desc())
    assert not _lib.hashgrid_backward_adam_supported(desc(dt=1))  # f16 table: the loss-scaled path
    assert not _lib.hashgrid_backward_adam_supported(desc(F=4))
    adam = lambda n, flags=0: _lib.adam_descriptor(n, 1e-2, 0.9, 0.99, 1e-15, 1e-15, 0.9, 0.99, 0.0, flags=flags)
    n = lv.n_rows * 2
    with pytest.raises(_lib.NGPError, match="adam.n"):
        _lib.call_raw("hashgrid_encode_backward_adam", [0] * 6, _lib.hashgrid_adam_descriptor(desc(), adam(n - 1)))
    with pytest.raises(_lib.NGPError, match="loss scaling"):
        _lib.call_raw("hashgrid_encode_backward_adam", [0] * 6,
                      _lib.hashgrid_adam_descriptor(desc(), adam(n, _lib.LOSS_SCALED)))
    with pytest.raises(_lib.NGPError, match="NULL"):
        _lib.call_raw("hashgrid_encode_backward_adam", [0] * 6, _lib.hashgrid_adam_descriptor(desc(), adam(n)))
    assert _lib.last_error() == (0, "")
