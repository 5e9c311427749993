"""Checkpoints in the reference's format (flax msgpack of NeRFState; see
jaxngp_amd/checkpoint.py).  flax is absent, so the encoding is pinned against
msgpack directly: the ext-type layout of arrays and scalars, the chunked-array
map, and the mapping between the flax kernels and the flat MLP vector."""
import numpy as np
import msgpack
import pytest
import torch

from jaxngp_amd import checkpoint as ck
from jaxngp_amd.nerfs import MLP_LAYERS, N_PARAMS


def test_array_ext_layout():
    """An ndarray is ext type 1 over msgpack [shape, dtype name, C bytes]; a numpy
    scalar is ext type 3 over the same layout of a 0-d array."""
    a = np.arange(6, dtype=np.float32).reshape(2, 3)
    raw = msgpack.unpackb(ck.to_bytes({"a": a, "s": np.int32(7)}), raw=False, strict_map_key=False)
    ea, es = raw["a"], raw["s"]
    assert isinstance(ea, msgpack.ExtType) and ea.code == 1
    shape, name, buf = msgpack.unpackb(ea.data, raw=True)
    assert shape == [2, 3] and name == b"float32" and buf == a.tobytes()
    assert isinstance(es, msgpack.ExtType) and es.code == 3
    shape, name, buf = msgpack.unpackb(es.data, raw=True)
    assert shape == [] and name == b"int32" and np.frombuffer(buf, np.int32)[0] == 7
    # a file laid out by hand with msgpack alone decodes to the same values
    hand = msgpack.packb({"a": msgpack.ExtType(1, msgpack.packb(((2, 3), "float32", a.tobytes()), use_bin_type=True)),
                          "n": None, "i": 5}, use_bin_type=True)
    got = ck.from_bytes(hand)
    np.testing.assert_array_equal(got["a"], a)
    assert got["n"] is None and got["i"] == 5


def test_round_trip_dtypes_and_chunking():
    rng = np.random.default_rng(0)
    state = {"f32": rng.normal(size=(5, 7)).astype(np.float32), "f16": rng.normal(size=9).astype(np.float16),
             "u8": rng.integers(0, 256, 17).astype(np.uint8), "b": rng.integers(0, 2, 33).astype(bool),
             "u32": rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32), "zero": np.zeros((3, 0), np.float32),
             "nested": {"x": np.int64(3), "none": None, "i": 2}}
    for max_chunk in (ck.MAX_CHUNK_BYTES, 16):
        data = ck.to_bytes(state, max_chunk)
        if max_chunk == 16:
            raw = msgpack.unpackb(data, raw=False, strict_map_key=False)
            assert raw["f32"][ck.CHUNKED] is True and raw["f32"]["shape"] == [5, 7]
        got = ck.from_bytes(data)
        for k in ("f32", "f16", "u8", "b", "u32", "zero"):
            assert got[k].dtype == state[k].dtype and got[k].shape == state[k].shape
            np.testing.assert_array_equal(got[k], state[k])
        assert got["nested"] == {"x": 3, "none": None, "i": 2}


def test_nerf_state_layout_and_mapping():
    """The NeRFState tree (utils/types.py:1127-1145, params from train.py:180-196) and
    the kernel mapping: kernel[i, j] of each Dense layer is flat[offset + i * out + j]
    in MLP_LAYERS order (nerfs.py:422-454 shapes)."""
    rng = np.random.default_rng(1)
    flat = rng.normal(size=N_PARAMS).astype(np.float32)
    table = rng.normal(size=(100, 2)).astype(np.float32)
    G3 = 16 ** 3
    dens = rng.uniform(size=G3).astype(np.float32)
    occ = dens > 0.5
    bits = np.packbits(occ.reshape(-1, 8), axis=-1, bitorder="little").reshape(-1)
    alive = np.arange(G3, dtype=np.int64)[::3]
    st = ck.nerf_state_dict(1234, table, flat, dens, occ, bits, alive, n_frames=7)
    assert set(st) == {"step", "params", "opt_state", "ogrid"}
    nerf = st["params"]["nerf"]
    assert set(nerf) == {"position_encoder", "density_mlp", "rgb_mlp"}
    assert nerf["position_encoder"][ck.TABLE_NAME].shape == (100, 2)
    assert st["params"]["appearance_embeddings"].shape == (7, 0) and st["params"]["bg"] is None
    o = 0
    for mlp, dense, fi, fo in MLP_LAYERS:
        k = nerf[mlp][dense]["kernel"]
        assert k.shape == (fi, fo)
        i, j = fi - 1, fo // 2
        assert k[i, j] == flat[o + i * fo + j]
        o += fi * fo
    got = ck.parse_nerf_state(ck.from_bytes(ck.to_bytes(st)))
    assert got["step"] == 1234
    np.testing.assert_array_equal(got["table"], table)
    np.testing.assert_array_equal(got["mlp"], flat)
    np.testing.assert_array_equal(got["occupancy"], bits)
    assert got["alive_indices"].dtype == np.uint32
    np.testing.assert_array_equal(got["alive_indices"], alive)
    # a step stored as flax does after training (a 0-d int32 array) reads the same
    st["step"] = np.asarray(1234, np.int32)
    assert ck.parse_nerf_state(ck.from_bytes(ck.to_bytes(st)))["step"] == 1234
    # wrong kernel shapes are refused
    nerf["rgb_mlp"]["Dense_2"]["kernel"] = np.zeros((3, 64), np.float32)
    with pytest.raises(ValueError):
        ck.parse_nerf_state(st)


def test_latest_checkpoint(tmp_path):
    for s in (5, 40, 300):
        (tmp_path / f"checkpoint_{s}").write_bytes(b"x")
    (tmp_path / "checkpoint_7.tmp").write_bytes(b"x")
    assert ck.latest_checkpoint(str(tmp_path)).endswith("checkpoint_300")
    assert ck.checkpoint_path(str(tmp_path), 12).endswith("checkpoint_12")


def test_background_model_parameters_round_trip(tmp_path):
    """params["bg"] holds a SkySphereBg's tree (rgb_mlp/Dense_0..2 kernels, [in, out])
    and restore_background loads it back."""
    from jaxngp_amd.background import SkySphereBg
    from jaxngp_amd.checkpoint import restore_background
    bg = SkySphereBg(4.0, n_extra_dims=2, seed=1)
    tree = bg.param_tree()
    assert [tree["rgb_mlp"][k]["kernel"].shape for k in ("Dense_0", "Dense_1", "Dense_2")] == [(22, 32), (32, 32),
                                                                                                (32, 3)]
    z = np.zeros(8, np.float32)
    state = ck.nerf_state_dict(3, np.zeros((8, 2), np.float32), np.zeros(9408, np.float32), z, z > 0,
                               np.zeros(1, np.uint8), np.zeros(0, np.uint32), bg_tree=tree)
    path = tmp_path / "checkpoint_3"
    path.write_bytes(ck.to_bytes(state))
    got = ck.load_checkpoint(str(path))
    np.testing.assert_array_equal(got["bg"]["rgb_mlp"]["Dense_1"]["kernel"], tree["rgb_mlp"]["Dense_1"]["kernel"])
    other = SkySphereBg(4.0, n_extra_dims=2, seed=2)
    assert not torch.equal(other.params, bg.params)
    restore_background(other, str(path))
    assert torch.equal(other.params, bg.params)
    # without a background model the field is nil, as the reference writes it
    state["params"]["bg"] = None
    assert ck.parse_nerf_state(ck.from_bytes(ck.to_bytes(state)))["bg"] is None
