"""Checkpoints in the reference's format: flax's msgpack serialization of NeRFState.

The reference saves its training state with flax.training.checkpoints
(app/nerf/train.py:200-206 restore, 252-261 save) and test.py restores it into an
empty NeRFState to render (app/nerf/test.py:53-72).  The state is a flax struct
(utils/types.py:1127-1145 NeRFState over flax's TrainState, 93-145
OccupancyDensityGrid); only its pytree fields are serialized.  The bytes are
flax.serialization's msgpack encoding (flax of the nixpkgs 22.11 pin in flake.nix;
flax is not in this image), restated here:

* the state is a msgpack map of field name -> value, nested dicts as maps with
  string keys, None as nil, Python ints as msgpack ints;
* a numpy array is msgpack ext type 1 whose data is the msgpack array
  [shape, dtype name, C-order bytes]; a numpy scalar is ext type 3 with the same
  data for a 0-d array;
* an array larger than 2^30 bytes is replaced by the map
  {"__msgpack_chunked_array__": True, "shape": shape, "chunks": {"0": ..., ...}}
  of consecutive pieces of its flattened data.

Parity against a file written by flax is unpinned: no checkpoint ships with the
reference and flax is absent here; the tests pin the encoding's layout (built with
msgpack alone) and the parameter mapping.

State dict, as written and as read:

    step:      int (a 0-d int32 array in files flax wrote after training)
    params:    {"nerf": {"position_encoder": {"latent codes stored on grid vertices": f32[rows, F]},
                         "density_mlp": {"Dense_0": {"kernel": f32[32, 64]}, "Dense_1": {"kernel": f32[64, 16]}},
                         "rgb_mlp": {"Dense_0": {"kernel": f32[32, 64]}, "Dense_1": {"kernel": f32[64, 64]},
                                     "Dense_2": {"kernel": f32[64, 3]}}},
                "bg": None, or {"rgb_mlp": {"Dense_0": {"kernel": f32[20+E, 32]}, "Dense_1": {"kernel":
                      f32[32, 32]}, "Dense_2": {"kernel": f32[32, 3]}}} with a sky-sphere background,
                "appearance_embeddings": f32[n_frames, 0]}
    opt_state: None when written (test.py restores into an empty state, whose
               opt_state is None, so rendering needs no optimizer state); ignored
               when read, so a restored Trainer starts its Adam moments at zero
    ogrid:     {"density": f32[K G^3], "occ_mask": bool[K G^3], "occupancy": u8[K G^3 / 8],
                "alive_indices": u32[n_alive]}
"""
from __future__ import annotations

import os
import re

import msgpack
import numpy as np

EXT_NDARRAY = 1
EXT_NPSCALAR = 3
MAX_CHUNK_BYTES = 2 ** 30
CHUNKED = "__msgpack_chunked_array__"
TABLE_NAME = "latent codes stored on grid vertices"


# ---------------------------------------------------------------------------
# encoding
# ---------------------------------------------------------------------------

def _array_bytes(a: np.ndarray) -> bytes:
    a = np.asarray(a, order="C")  # keeps 0-d arrays 0-d
    return msgpack.packb((list(a.shape), a.dtype.name, a.tobytes("C")), use_bin_type=True)


def _array_from(data: bytes) -> np.ndarray:
    shape, dtype_name, buf = msgpack.unpackb(data, raw=True)
    name = dtype_name.decode() if isinstance(dtype_name, bytes) else dtype_name
    return np.frombuffer(buf, dtype=np.dtype(name)).reshape(tuple(shape), order="C").copy()


def _ext_pack(x):
    if isinstance(x, np.ndarray):
        return msgpack.ExtType(EXT_NDARRAY, _array_bytes(x))
    if isinstance(x, np.generic):
        return msgpack.ExtType(EXT_NPSCALAR, _array_bytes(np.asarray(x)))
    raise TypeError(f"cannot serialize {type(x).__name__}")


def _ext_unpack(code: int, data: bytes):
    if code == EXT_NDARRAY:
        return _array_from(data)
    if code == EXT_NPSCALAR:
        return _array_from(data)[()]
    return msgpack.ExtType(code, data)


def _chunk(tree, max_bytes: int):
    if isinstance(tree, dict):
        return {k: _chunk(v, max_bytes) for k, v in tree.items()}
    if isinstance(tree, np.ndarray) and tree.nbytes > max_bytes:
        flat = np.ascontiguousarray(tree).reshape(-1)
        per = max(1, max_bytes // max(1, flat.itemsize))
        chunks = {str(i): flat[o:o + per] for i, o in enumerate(range(0, flat.size, per))}
        return {CHUNKED: True, "shape": list(tree.shape), "chunks": chunks}
    return tree


def _unchunk(tree):
    if isinstance(tree, dict):
        if tree.get(CHUNKED) is True:
            chunks = tree["chunks"]
            flat = np.concatenate([chunks[str(i)] for i in range(len(chunks))])
            return flat.reshape(tuple(tree["shape"]))
        return {k: _unchunk(v) for k, v in tree.items()}
    return tree


def to_bytes(state: dict, max_chunk_bytes: int = MAX_CHUNK_BYTES) -> bytes:
    """msgpack bytes of a state dict (nested dicts of numpy arrays, ints, None)."""
    return msgpack.packb(_chunk(state, max_chunk_bytes), default=_ext_pack, strict_types=True, use_bin_type=True)


def from_bytes(data: bytes) -> dict:
    return _unchunk(msgpack.unpackb(data, ext_hook=_ext_unpack, raw=False, strict_map_key=False))


# ---------------------------------------------------------------------------
# NeRF state <-> state dict
# ---------------------------------------------------------------------------

def _mlp_tree(flat: np.ndarray) -> dict:
    from .nerfs import MLP_LAYERS
    tree, o = {}, 0
    for mlp, dense, fi, fo in MLP_LAYERS:
        tree.setdefault(mlp, {})[dense] = {"kernel": np.asarray(flat[o:o + fi * fo], np.float32).reshape(fi, fo)}
        o += fi * fo
    return tree


def _mlp_flat(tree: dict) -> np.ndarray:
    from .nerfs import MLP_LAYERS
    parts = []
    for mlp, dense, fi, fo in MLP_LAYERS:
        k = np.asarray(tree[mlp][dense]["kernel"], np.float32)
        if k.shape != (fi, fo):
            raise ValueError(f"{mlp}/{dense}/kernel has shape {k.shape}, expected {(fi, fo)}")
        parts.append(k.reshape(-1))
    return np.concatenate(parts)


def nerf_state_dict(step: int, table: np.ndarray, mlp_flat: np.ndarray, density: np.ndarray, occ_mask: np.ndarray,
                    occupancy: np.ndarray, alive_indices: np.ndarray, n_frames: int = 0,
                    bg_tree: dict | None = None) -> dict:
    """The state dict of a NeRFState (utils/types.py:1127-1145) with the given arrays;
    `bg_tree` is a SkySphereBg's parameter tree (params["bg"], nerfs.py:133-213)."""
    nerf = {"position_encoder": {TABLE_NAME: np.asarray(table, np.float32)}}
    nerf.update(_mlp_tree(np.asarray(mlp_flat, np.float32)))
    return {
        "step": int(step),
        "params": {"nerf": nerf, "bg": bg_tree,
                   "appearance_embeddings": np.zeros((int(n_frames), 0), np.float32)},
        "opt_state": None,
        "ogrid": {"density": np.asarray(density, np.float32), "occ_mask": np.asarray(occ_mask, np.bool_),
                  "occupancy": np.asarray(occupancy, np.uint8),
                  "alive_indices": np.asarray(alive_indices, np.uint32)},
    }


def parse_nerf_state(state: dict) -> dict:
    """{"step", "table", "mlp", "density", "occ_mask", "occupancy", "alive_indices", "bg"}
    from a NeRFState state dict (either writer's); "bg" is the background model's
    parameter tree or None."""
    step = state.get("step")
    step = int(np.asarray(step)) if step is not None else 0
    nerf = state["params"]["nerf"]
    out = {"step": step, "table": np.asarray(nerf["position_encoder"][TABLE_NAME], np.float32),
           "mlp": _mlp_flat(nerf), "bg": state["params"].get("bg")}
    og = state.get("ogrid") or {}
    for k in ("density", "occ_mask", "occupancy", "alive_indices"):
        out[k] = None if og.get(k) is None else np.asarray(og[k])
    return out


# ---------------------------------------------------------------------------
# files and the Trainer
# ---------------------------------------------------------------------------

def checkpoint_path(ckpt_dir: str, step: int) -> str:
    """flax's file name for a step: <dir>/checkpoint_<step>."""
    return os.path.join(ckpt_dir, f"checkpoint_{int(step)}")


def latest_checkpoint(path: str) -> str:
    """`path` itself if it is a file, else the checkpoint_<n> in it with the largest n."""
    if os.path.isfile(path):
        return path
    best = None
    for name in os.listdir(path):
        m = re.fullmatch(r"checkpoint_(\d+)", name)
        if m and (best is None or int(m.group(1)) > best[0]):
            best = (int(m.group(1)), name)
    if best is None:
        raise FileNotFoundError(f"no checkpoint_<step> file in {path!r}")
    return os.path.join(path, best[1])


def save_checkpoint(ckpt_dir: str, trainer, n_frames: int | None = None, bg_model=None) -> str:
    """Writes the trainer's field and occupancy grid as <dir>/checkpoint_<step> in the
    reference's format and returns the path.  The table and MLP are f32 (the master
    weights on the f16 path); `bg_model` (a SkySphereBg) goes to params["bg"]."""
    import torch

    os.makedirs(ckpt_dir, exist_ok=True)
    if hasattr(trainer, "sync"):
        trainer.sync()  # the last update may still run on the trainer's side stream
    if bg_model is None:
        bg_model = getattr(trainer, "bg_model", None)
    # the f16 path keeps f32 master weights beside its f16 table; the f32 path updates
    # the model's table in place
    table = trainer.table_master if getattr(trainer, "fp16", False) \
        else trainer.model.position_encoder.table.detach().float()
    g = trainer.grid
    if n_frames is None:
        n_frames = int(trainer.data.rgba.shape[0]) if hasattr(trainer, "data") else 0
    with torch.no_grad():
        state = nerf_state_dict(trainer.step_count, table.cpu().numpy(), trainer.model.mlp_params.detach().cpu().numpy(),
                                g.density.cpu().numpy(), g.occ_mask.cpu().numpy(), g.bitfield.cpu().numpy(),
                                g.alive.cpu().numpy(), n_frames,
                                bg_model.param_tree() if bg_model is not None else None)
    path = checkpoint_path(ckpt_dir, trainer.step_count)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(to_bytes(state))
    os.replace(tmp, path)
    return path


def load_checkpoint(path: str) -> dict:
    """Parsed NeRF state (see parse_nerf_state) of a checkpoint file or the latest in a directory."""
    with open(latest_checkpoint(path), "rb") as f:
        return parse_nerf_state(from_bytes(f.read()))


def restore_trainer(trainer, path: str) -> int:
    """Loads a checkpoint into a Trainer built for the same scene options: field
    weights (and their f16 copies on the precision path), the background model's
    parameters when the trainer has one, occupancy grid and step.
    The Adam moments restart at zero (the file carries no optimizer state).
    Returns the step."""
    import torch

    ck = load_checkpoint(path)
    if hasattr(trainer, "sync"):
        trainer.sync()  # no pending update may overwrite what is restored
    m = trainer.model
    dev = m.mlp_params.device
    table = m.position_encoder.table
    if tuple(ck["table"].shape) != tuple(table.shape):
        raise ValueError(f"table shape {ck['table'].shape} does not match the model's {tuple(table.shape)}")
    with torch.no_grad():
        t = torch.from_numpy(ck["table"]).to(dev)
        table.copy_(t.to(table.dtype))
        if getattr(trainer, "table_master", None) is not None:
            trainer.table_master.copy_(t)
        m.mlp_params.copy_(torch.from_numpy(ck["mlp"]).to(dev))
        if getattr(m, "mlp_half", None) is not None:
            m.mlp_half.copy_(m.mlp_params.half())
        g = trainer.grid
        if ck["density"] is not None:
            g.density.copy_(torch.from_numpy(ck["density"]).to(dev))
            g.occ_mask.copy_(torch.from_numpy(ck["occ_mask"]).to(dev))
            g.bitfield.copy_(torch.from_numpy(ck["occupancy"]).to(dev))
            alive = ck["alive_indices"].astype(np.int64)
            g.alive = torch.from_numpy(alive).to(dev)
            # alive_indices_offset is not a pytree field (types.py:107): rebuilt per cascade
            G3 = g.G ** 3
            counts = [int(((alive >= c * G3) & (alive < (c + 1) * G3)).sum()) for c in range(g.K)]
            g.alive_offsets = np.cumsum([0] + counts).tolist()
        for v in trainer.opt.values():
            v.zero_()
        if getattr(trainer, "bg_model", None) is not None:
            if ck["bg"] is None:
                raise ValueError("the trainer has a background model but the checkpoint holds none")
            trainer.bg_model.load_param_tree(ck["bg"])
            trainer.bb["mm"].zero_()
            trainer.bb["mv"].zero_()
    trainer.step_count = ck["step"]
    trainer._calls = None  # rebind the prepared calls to the restored buffers
    return ck["step"]


def restore_background(bg_model, path: str) -> None:
    """Loads params["bg"] of a checkpoint into a SkySphereBg of the same width."""
    ck = load_checkpoint(path)
    if ck["bg"] is None:
        raise ValueError(f"{path!r} holds no background model parameters")
    bg_model.load_param_tree(ck["bg"])
