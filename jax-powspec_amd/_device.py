"""Torch plumbing: device tensors, stream handles, per-device plans and workspaces.

PyTorch only provides device memory and streams here; all arithmetic of the path runs in
libpowspec_amd.so.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np
import torch

from . import _lib

_plans: dict = {}
_plans_lock = threading.Lock()


def require_gpu(device=None) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("jax_powspec_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError(f"expected a cuda (ROCm) device, got {device}")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return device


def stream_ptr(device: torch.device, stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream(device)
    return int(s.cuda_stream)


def ptr(t) -> int | None:
    return None if t is None else int(t.data_ptr())


class Plan:
    """Owns a pk_plan (twiddle + window tables) for one (device, n)."""

    def __init__(self, n: int, device: torch.device):
        lib = _lib.load()
        self.n = int(n)
        self.device = device
        h = C.c_void_p()
        with torch.cuda.device(device):
            _lib.check("pk_plan_create", lib.pk_plan_create(self.n, C.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if getattr(self, "handle", None) and self.handle.value:
                with torch.cuda.device(self.device):
                    _lib.load().pk_plan_destroy(self.handle)
        except Exception:
            pass


def plan(n: int, device: torch.device) -> Plan:
    key = (device.index, int(n))
    with _plans_lock:
        p = _plans.get(key)
        if p is None:
            if not _lib.load().pk_supported_n(int(n)):
                raise ValueError(f"unsupported mesh side n={n}: need even 8 <= n <= 1024 of the form 2^a or 3*2^a")
            p = Plan(n, device)
            _plans[key] = p
        return p


def workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def as_f32_vector(a, device: torch.device, name: str) -> tuple[torch.Tensor, int]:
    """1-D float32 device view (no copy when already one) and its element stride."""
    if isinstance(a, torch.Tensor):
        t = a.detach()
        if t.device != device:
            t = t.to(device)
        if t.dtype != torch.float32:
            t = t.to(torch.float32)
    else:
        t = torch.as_tensor(np.asarray(a, dtype=np.float32), device=device)
    if t.dim() != 1:
        t = t.reshape(-1)
    st = t.stride(0) if t.numel() > 1 else 1
    if st < 1:
        t = t.contiguous()
        st = 1
    return t, int(st)


def as_f32_scalar(v) -> float:
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().item()
    return float(np.float32(v))


def as_mesh(delta, device: torch.device, name: str = "delta") -> torch.Tensor:
    if isinstance(delta, torch.Tensor):
        t = delta.detach()
        if t.device != device:
            t = t.to(device)
    else:
        t = torch.as_tensor(np.asarray(delta), device=device)
    if t.dim() != 3 or not (t.shape[0] == t.shape[1] == t.shape[2]):
        raise ValueError(f"{name} must be a cubic (n, n, n) mesh, got shape {tuple(t.shape)}")
    if t.dtype != torch.float32:
        t = t.to(torch.float32)
    return t.contiguous()


def device_of(*objs):
    for o in objs:
        if isinstance(o, torch.Tensor) and o.is_cuda:
            return require_gpu(o.device)
    return require_gpu(None)
