"""Device buffers: PyTorch-ROCm tensors are used only as device memory and streams.

Unsigned index arrays are stored in signed tensors of the same width (int32 for
uint32, int16 for uint16) because the kernels only need the bits; ``to_host`` views
them back as unsigned numpy arrays.
"""
from __future__ import annotations

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("yasps_amd: no HIP device is visible (the solve path runs on the GPU only)")
    return t.device("cuda", t.cuda.current_device())


def ptr(t):
    """Device address of a tensor, or None for an empty / missing one."""
    if t is None or t.numel() == 0:
        return None
    return t.data_ptr()


def stream():
    return torch().cuda.current_stream().cuda_stream


_STORAGE = {"float64": "float64", "float32": "float32", "uint32": "int32", "int32": "int32",
            "uint16": "int16", "int16": "int16", "int64": "int64", "uint8": "uint8"}


def _tdtype(name):
    return getattr(torch(), _STORAGE[name])


def to_device(arr, dtype: str):
    a = np.ascontiguousarray(np.asarray(arr).astype(dtype, copy=False))
    if dtype == "uint32":
        a = a.view(np.int32)
    elif dtype == "uint16":
        a = a.view(np.int16)
    return torch().from_numpy(a.copy()).to(device())


_CONSTANTS = {}


def constant(values, dtype: str):
    """Device copy of a small host array that no one writes (a kernel's parameters, a
    child count), uploaded once per distinct content. An upload from pageable host memory
    synchronises the stream, so one per launch would stall the device between launches."""
    a = np.ascontiguousarray(np.asarray(values).astype(dtype, copy=False))
    key = (dtype, a.shape, a.tobytes())
    t = _CONSTANTS.get(key)
    if t is None:
        if len(_CONSTANTS) >= 4096:
            _CONSTANTS.clear()
        t = _CONSTANTS[key] = to_device(a, dtype)
    return t


def constant_full(n: int, value: float, dtype: str = "float64"):
    """``constant`` of n copies of one value (keyed by n and the value)."""
    key = (dtype, "full", int(n), float(value))
    t = _CONSTANTS.get(key)
    if t is None:
        if len(_CONSTANTS) >= 4096:
            _CONSTANTS.clear()
        t = _CONSTANTS[key] = to_device(np.full(int(n), float(value)), dtype)
    return t


def empty(n, dtype: str):
    return torch().empty(int(n), dtype=_tdtype(dtype), device=device())


def zeros(n, dtype: str):
    return torch().zeros(int(n), dtype=_tdtype(dtype), device=device())


def to_host(t, dtype: str | None = None):
    out = t.detach().cpu().numpy()
    if dtype == "uint32":
        out = out.view(np.uint32)
    elif dtype == "uint16":
        out = out.view(np.uint16)
    elif dtype is not None:
        out = out.astype(dtype)
    return out
