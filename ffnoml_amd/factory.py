"""Model construction and checkpoint IO with the reference's interfaces.

create_ffno3d mirrors ffno_model_factory.py:6-19 (the plugin the service resolves by factory_module/factory_name,
ffno_fsdp_runtime.py:145-153); the checkpoint dict is the reference's (train_utils.py:118-138: model_state,
normalization_stats {mean_X, std_X, mean_Y, std_Y}, io_metadata {Cin, Cout}), loaded only with weights_only=True.
"""
import os

import numpy as np
import torch

from .atoms import ATOMS, MODEL_CONFIG, species as species_of
from .model import FFNO3D


def create_ffno3d(Cin, Cout, device="cpu", atoms=None, **overrides):
    """FFNO3D from the final model's configuration plus overrides; atoms (e.g. ["CA"]) selects the head's species,
    otherwise FFNOML_ACTIVE_ATOMS (default "H") as the reference's config does."""
    configuration = dict(MODEL_CONFIG)
    configuration.update(overrides)
    for key in ("spec_dropout", "vertical_dropout", "spec_dropout_layers", "vertical_dropout_layers", "dropout"):
        configuration.pop(key, None)          # inactive in eval mode; the engine has no dropout
    names = atoms if atoms is not None else os.environ.get("FFNOML_ACTIVE_ATOMS", "H").split(",")
    if "species" not in overrides or atoms is not None:
        configuration["species"] = species_of(names)
    configuration["in_channels"] = int(Cin)
    configuration["out_channels"] = int(Cout)
    model = FFNO3D(**configuration)
    return model.to(device) if str(device) != "cpu" else model


def load_checkpoint(path, map_location="cpu"):
    """torch.load with weights_only=True; numpy arrays in normalization_stats are allowed through the safe-globals
    list (train_utils.py:20-91 without its unsafe fallback)."""
    try:
        return torch.load(path, map_location=map_location, weights_only=True)
    except Exception as first:
        allowed = [x for x in (getattr(np.core.multiarray, "_reconstruct", None), np.ndarray, np.dtype,
                               getattr(np.core.multiarray, "scalar", None)) if x is not None]
        try:
            allowed.append(type(np.dtype(np.float32)))
        except Exception:
            pass
        try:
            with torch.serialization.safe_globals(allowed):
                return torch.load(path, map_location=map_location, weights_only=True)
        except Exception:
            raise first


def get_checkpoint_normalization(ckpt):
    stats = ckpt.get("normalization_stats") if isinstance(ckpt, dict) else None
    if not isinstance(stats, dict) or not all(k in stats for k in ("mean_X", "std_X", "mean_Y", "std_Y")):
        return None
    return {k: np.asarray(torch.as_tensor(stats[k], dtype=torch.float32).cpu().numpy()) for k in
            ("mean_X", "std_X", "mean_Y", "std_Y")}


def get_checkpoint_io_metadata(ckpt):
    meta = ckpt.get("io_metadata") if isinstance(ckpt, dict) else None
    if not isinstance(meta, dict) or "Cin" not in meta or "Cout" not in meta:
        return None
    return {"Cin": int(meta["Cin"]), "Cout": int(meta["Cout"])}


def save_checkpoint(path, model, normalization=None, io_metadata=None, **extra):
    """The reference's checkpoint dict, written atomically through a .tmp file (train_utils.py:134-138)."""
    ckpt = dict(extra)
    ckpt["model_state"] = {k: v.detach().cpu() for k, v in model.state_dict().items()}
    if normalization is not None:
        ckpt["normalization_stats"] = {k: np.asarray(v, np.float32) for k, v in normalization.items()}
    if io_metadata is not None:
        ckpt["io_metadata"] = {"Cin": int(io_metadata["Cin"]), "Cout": int(io_metadata["Cout"])}
    tmp = f"{path}.tmp"
    torch.save(ckpt, tmp)
    os.replace(tmp, path)
