"""Mass assignment -- drop-in for the reference's src/mas.py, on ROCm torch tensors.

``cic_mas_vec`` and ``cic_mas`` keep the reference signatures and results (including the
(ix, iyp, izp) corner weight of src/mas.py:149 / :76, the truncating cell index and the
reference's wrap rules).  ``assign`` exposes the extensions: textbook mass-conserving
CIC, TSC, tile shape selection and the deposited-mass statistics.

All arithmetic runs in libpowspec_amd.so (pk_scatter); this module only prepares device
pointers.  Like the reference the functions are functional: ``delta`` is the
accumulation base and is not modified.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _device, _lib

__all__ = ["cic_mas_vec", "cic_mas", "tsc_mas_vec", "assign"]


def _scatter(delta, x, y, z, w, xmin, ymin, zmin, box_size, n_bins, stencil, wrap_mode, tile=None,
             stats=None, out=None):
    n = int(n_bins)
    device = _device.device_of(delta, x, y, z, w)
    xs, sx = _device.as_f32_vector(x, device, "x")
    ys, sy = _device.as_f32_vector(y, device, "y")
    zs, sz = _device.as_f32_vector(z, device, "z")
    np_ = xs.numel()
    if ys.numel() != np_ or zs.numel() != np_:
        raise ValueError("x, y, z must have the same length")
    ws_, sw = (None, 1)
    if w is not None:
        ws_, sw = _device.as_f32_vector(w, device, "w")
        if ws_.numel() != np_:
            raise ValueError("w must have the same length as x")
    mesh_in = None
    if delta is not None:
        mesh_in = _device.as_mesh(delta, device)
        if mesh_in.shape[0] != n:
            raise ValueError(f"delta has side {mesh_in.shape[0]} but n_bins={n}")
    if out is None:
        out = torch.empty((n, n, n), dtype=torch.float32, device=device)
    tptr, _keep = _lib.tile_arg(tile)
    lib = _lib.load()
    nbytes = C.c_size_t()
    _lib.check("pk_scatter_workspace_size",
               lib.pk_scatter_workspace_size(n, np_, _lib.STENCIL[stencil], tptr, C.byref(nbytes)))
    with torch.cuda.device(device):
        ws = _device.workspace(nbytes.value, device)
        _lib.check("pk_scatter", lib.pk_scatter(
            _device.ptr(xs), sx, _device.ptr(ys), sy, _device.ptr(zs), sz, _device.ptr(ws_), sw, np_,
            _device.as_f32_scalar(xmin), _device.as_f32_scalar(ymin), _device.as_f32_scalar(zmin),
            float(box_size), n, _lib.STENCIL[stencil], int(wrap_mode), tptr,
            _device.ptr(mesh_in), _device.ptr(out), _device.ptr(stats), _device.ptr(ws), nbytes.value,
            _device.stream_ptr(device)))
    return out


def cic_mas_vec(delta, x, y, z, w, n_part, xmin, ymin, zmin, box_size, n_bins, wrap):
    """Vectorised CIC of the reference, src/mas.py:88-153 (``n_part`` is unused, as there).

    Returns a new (n_bins,)*3 float32 tensor = delta + deposits.  wrap=True uses
    ixp = (ix + 1 + n) % n (:119-122); wrap=False the reference's false_wrap (:123-134),
    whose ddx test never fires.  Out-of-range base indices follow JAX scatter rules
    (negative -> +n, still outside [0, n) -> dropped).
    """
    mode = _lib.WRAP_VEC_TRUE if bool(wrap) else _lib.WRAP_VEC_FALSE
    return _scatter(delta, x, y, z, w, xmin, ymin, zmin, box_size, n_bins, "CIC_REF", mode)


def cic_mas(delta, x, y, z, w, n_part, xmin, ymin, zmin, box_size, n_bins, wrap):
    """Scan-based CIC of the reference, src/mas.py:5-87, computed in one parallel pass.

    wrap=True: ixp -> ixp - n when ixp >= n (:20-28); wrap=False: ixp, ddx -> 0 when
    ixp >= n while mdx keeps 1 - ddx (:30-38, :56-58).
    """
    mode = _lib.WRAP_SCAN_TRUE if bool(wrap) else _lib.WRAP_SCAN_FALSE
    return _scatter(delta, x, y, z, w, xmin, ymin, zmin, box_size, n_bins, "CIC_REF", mode)


def tsc_mas_vec(delta, x, y, z, w, n_part, xmin, ymin, zmin, box_size, n_bins, wrap=True):
    """Periodic TSC with the reference's signature (extension; no reference counterpart).

    Nearest grid point i0 = floor(pos + 1/2), d = pos - i0, weights 1/2 (1/2 - d)^2,
    3/4 - d^2, 1/2 (1/2 + d)^2 on i0 - 1, i0, i0 + 1 (periodic).
    """
    return _scatter(delta, x, y, z, w, xmin, ymin, zmin, box_size, n_bins, "TSC", 0)


def assign(x, y, z, w=None, *, box_size, n_bins, mas="CIC_REF", wrap=True, xmin=0.0, ymin=0.0, zmin=0.0,
           delta=None, tile=None, stats=None, out=None):
    """Extended mass assignment.

    mas: 'CIC_REF' (reference stencil, cic_mas_vec), 'CIC' (textbook, mass-conserving,
    periodic) or 'TSC'.  delta: optional accumulation base (None = zeros).
    tile: (1, ty, 0) row-bundle tile of the sorted deposit (ty rows of full z extent, ty | n),
          'direct' for global atomics, or None (ty = largest divisor of n <= 32).
    stats: optional float64 device tensor of 4 entries receiving [mass, n^3/mass, -, -].
    """
    mas = mas.upper()
    if mas not in _lib.STENCIL:
        raise ValueError(f"unknown mass-assignment scheme {mas!r}")
    mode = _lib.WRAP_VEC_TRUE if bool(wrap) else _lib.WRAP_VEC_FALSE
    return _scatter(delta, x, y, z, w, xmin, ymin, zmin, box_size, n_bins, mas, mode, tile=tile, stats=stats,
                    out=out)
