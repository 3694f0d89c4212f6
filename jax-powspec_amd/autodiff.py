"""Differentiable P_l(k): torch.autograd around pk_estimate / pk_estimate_backward.

The reference differentiates its JAX path with jax.value_and_grad (e.g. a loss on P(k)
with respect to particle positions, tests/lognormal.py:101, or to weights,
tests/bias.py:66).  Here the backward is hand written: the adjoint of the binning (a
forward FFT of the saved spectrum, the per-mode chain rule, inverse FFTs back to the mesh)
followed by a gather that transposes the scatter, including the dependence of the mean
density on the particles.

    pos.requires_grad_()
    k, pk, nmodes = estimate_ad(pos, box_size=L, n_bins=n, k_edges=edges, mas="CIC")
    loss = (((pk[:, 0] - target) / sigma) ** 2).sum()
    loss.backward()            # pos.grad: dLoss/dpos, shape (N, 3)

Only the multipoles carry gradients; mode counts and k are constants.  The forward
workspace (spectrum after the y pass) is kept until the backward has run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _device, _lib, binning
from .correlations import _Outputs, _window_p

__all__ = ["estimate_ad"]


class _Estimate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, w, cfg):
        device = pos.device
        n = cfg["n"]
        plan = _device.plan(n, device)
        pos_c = pos.detach().to(torch.float32).contiguous()
        w_c = None if w is None else w.detach().to(torch.float32).contiguous()
        npart = pos_c.shape[0]
        tab = cfg["tables"]
        nb = tab.nb
        out = _Outputs(nb, device, bool(cfg["flags"] & _lib.FLAG_KMEAN))
        stats = torch.zeros(_lib.STATS_LEN, dtype=torch.float64, device=device)
        lib = _lib.load()
        tptr, tkeep = _lib.tile_arg(cfg["tile"])
        nbytes = C.c_size_t()
        base = _device.ptr(pos_c)
        with torch.cuda.device(device):
            _lib.check("pk_estimate_workspace_size", lib.pk_estimate_workspace_size(
                plan.handle, npart, cfg["stencil"], tptr, nb, C.byref(nbytes)))
            ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=device)
            _lib.check("pk_estimate", lib.pk_estimate(
                plan.handle, base, 3, base + 4 if base else None, 3, base + 8 if base else None, 3, _device.ptr(w_c), 1,
                npart, 0.0, 0.0, 0.0, cfg["box"], cfg["stencil"], cfg["wrap"], tptr, _device.ptr(tab.thr),
                _device.ptr(tab.jst), _device.ptr(tab.modes), _device.ptr(tab.ksum), nb, cfg["p"], cfg["flags"], cfg["shot"], _device.ptr(out.sums),
                _device.ptr(out.counts), _device.ptr(out.pk), _device.ptr(out.nmodes), _device.ptr(out.kmean),
                _device.ptr(stats), _device.ptr(ws), nbytes.value, _device.stream_ptr(device)))
        ctx.cfg = cfg
        ctx.has_w = w is not None
        ctx.save_for_backward(pos_c, w_c if w_c is not None else torch.empty(0, device=device), ws, stats,
                              out.counts)
        kmean = out.kmean if out.kmean is not None else torch.empty(0, dtype=torch.float64, device=device)
        ctx.mark_non_differentiable(out.nmodes, kmean)
        return out.pk, out.nmodes, kmean

    @staticmethod
    def backward(ctx, grad_pk, grad_nmodes, grad_kmean):
        pos_c, w_c, ws, stats, counts = ctx.saved_tensors
        cfg = ctx.cfg
        device = pos_c.device
        n = cfg["n"]
        plan = _device.plan(n, device)
        tab = cfg["tables"]
        nb = tab.nb
        npart = pos_c.shape[0]
        if grad_pk is None:
            return None, None, None
        gpk = torch.nan_to_num(grad_pk.to(torch.float64), nan=0.0).contiguous()
        grad_pos = torch.empty((npart, 3), dtype=torch.float32, device=device)
        want_w = ctx.has_w and ctx.needs_input_grad[1]
        grad_w = torch.empty(npart, dtype=torch.float32, device=device) if want_w else None
        lib = _lib.load()
        tptr, tkeep = _lib.tile_arg(cfg["tile"])
        nbytes = C.c_size_t()
        base = _device.ptr(pos_c)
        with torch.cuda.device(device):
            _lib.check("pk_estimate_backward_workspace_size", lib.pk_estimate_backward_workspace_size(
                plan.handle, npart, nb, C.byref(nbytes)))
            bws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=device)
            _lib.check("pk_estimate_backward", lib.pk_estimate_backward(
                plan.handle, base, 3, base + 4 if base else None, 3, base + 8 if base else None, 3,
                _device.ptr(w_c) if ctx.has_w else None, 1, npart, 0.0, 0.0, 0.0, cfg["box"], cfg["stencil"],
                cfg["wrap"], tptr, _device.ptr(tab.thr), _device.ptr(tab.jst), nb, cfg["p"], _device.ptr(ws), ws.numel(),
                _device.ptr(stats), _device.ptr(gpk), _device.ptr(counts), _device.ptr(grad_pos),
                _device.ptr(grad_w), _device.ptr(bws), nbytes.value, _device.stream_ptr(device)))
        return grad_pos, grad_w, None


def estimate_ad(pos, w=None, *, box_size, n_bins, k_edges=None, mas="CIC_REF", wrap=True, window=None,
                shot_noise=0.0, tile=None):
    """Differentiable estimate(): P_l(k) of particles pos (N, 3) float32 on the GPU.

    Same arguments and results as correlations.estimate (positions with xmin = 0);
    Pk3D carries gradients to ``pos`` and, when given and requiring grad, to ``w``.
    Returns (k3D, Pk3D, Nmodes3D).
    """
    mas = mas.upper()
    if mas not in _lib.STENCIL:
        raise ValueError(f"unknown mass-assignment scheme {mas!r}")
    device = _device.require_gpu(pos.device)
    if pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError("pos must have shape (N, 3)")
    if w is not None and w.shape != (pos.shape[0],):
        raise ValueError("w must have shape (N,)")
    n = int(n_bins)
    fundamental = k_edges is None
    edges = binning.fundamental_grid_edges(n) if fundamental else binning.grid_edges(k_edges, box_size)
    tables = binning.TABLES.get(device, n, edges)
    p = _window_p(window if window is not None else ("TSC" if mas == "TSC" else "CIC"))
    cfg = dict(n=n, box=float(box_size), stencil=_lib.STENCIL[mas],
               wrap=_lib.WRAP_VEC_TRUE if wrap else _lib.WRAP_VEC_FALSE, p=p,
               flags=_lib.FLAG_KMEAN if fundamental else 0, shot=float(shot_noise), tile=tile, tables=tables)
    pk, nmodes, kmean = _Estimate.apply(pos, w, cfg)
    kF = 2.0 * np.pi / float(box_size)
    if fundamental:
        k3D = (kmean.detach() * kF).to(torch.float32)
    else:
        k3D = torch.as_tensor((0.5 * (edges[1:] + edges[:-1]) * kF).astype(np.float32), device=device)
    return k3D, pk, nmodes
