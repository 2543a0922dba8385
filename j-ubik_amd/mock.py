"""Mock data: Poisson counts drawn on the GPU from the expected counts.

jubik/data.py:81-151 (create_mock_data) evaluates lam = R(sky(position)),
read out per key, draws jax.random.poisson per key (data.py:139-145),
casting to int, and saves the data and the position as pickles of their
key -> array dicts (data.py:146-149, utils.py:35-47, 170-182).  Here the draw is libjubik_hip's counter-based Poisson
sampler: each count is a pure function of (seed, flat index, lam), so the
same seed reproduces the data bit for bit and a sub-array draws the same
counts as the full array.
"""
import os
import pickle

import numpy as np
import torch

from . import _device as dev
from . import _lib
from .data import Vector


def poisson_sample(lam, seed=0):
    """Poisson(lam) counts, elementwise; int32, same shape and kind as lam
    (a device tensor stays on the device)."""
    dev.require_device()
    if isinstance(lam, torch.Tensor) and lam.dtype in (torch.float32, torch.float64):
        lt = lam.to("cuda").contiguous()
    else:
        lt = dev.to_device(lam, torch.float64)
    if lt.numel() and float(lt.min()) < 0:
        raise ValueError("Poisson rates must be non-negative")
    out = torch.empty(lt.shape, dtype=torch.int32, device="cuda")
    _lib.call("jubik_poisson_sample", dev.code(lt.dtype), dev.ptr(lt), int(lt.numel()),
              int(seed) & ((1 << 64) - 1), dev.ptr(out), dev.stream())
    return dev.like_input(out, lam)


def create_mock_counts(lam, keep, seed=0, keys=None):
    """Read-out mock data of expected counts ``lam`` (M, E, H, W): a Vector
    {key_i: int64 counts of module i's kept pixels in C order}, the wire
    format of jubik/response.py:110 and data.py:142-145."""
    keep = np.asarray(keep, dtype=bool)
    lam_shape = tuple(lam.shape)
    if lam_shape != keep.shape:
        raise ValueError(f"lam {lam_shape} and keep {keep.shape} must have the same shape")
    keys = list(range(keep.shape[0])) if keys is None else list(keys)
    if len(keys) != keep.shape[0]:
        raise ValueError("length of keys should match the number of modules")
    counts = poisson_sample(lam, seed)
    counts = counts.cpu().numpy() if isinstance(counts, torch.Tensor) else counts
    return Vector({k: counts[i][keep[i]].astype(np.int64) for i, k in enumerate(keys)})


def _host_tree(tree):
    out = {}
    for k, v in dict(tree).items():
        out[k] = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
    return out


def save_mock_data(data, position, file_info):
    """Write mock data and the position it was drawn at as the reference
    does (data.py:146-149): the data Vector's key -> counts dict to
    file_info["res_dir"]/file_info["data_dict"] and the position's key ->
    array dict to file_info["res_dir"]/file_info["pos_dict"], each with
    pickle.dump, creating the directory if needed.  Returns both paths."""
    res_dir = file_info["res_dir"]
    os.makedirs(res_dir, exist_ok=True)
    paths = (os.path.join(res_dir, file_info["data_dict"]),
             os.path.join(res_dir, file_info["pos_dict"]))
    tree = data.tree if isinstance(data, Vector) else data
    pos = position.tree if isinstance(position, Vector) else position
    for obj, path in ((tree, paths[0]), (pos, paths[1])):
        with open(path, "wb") as f:
            pickle.dump(_host_tree(obj), f)
    return paths
