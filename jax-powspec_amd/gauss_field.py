"""Seeded lognormal mocks on the GPU -- drop-in for src/gauss_field.py and src/populate_field.py.

``gaussian_field`` (src/gauss_field.py:5-80) keeps the reference signature and returns the
half spectrum delta_k, complex64 of shape (grid, grid, grid//2+1) like the reference's
array (a view of the device layout [kz][kx][ky]).  ``irfftn`` is the inverse transform
the reference's caller applies (scipy.fft.irfftn, tests/create_lognormal.py:47).
``populate_field`` (src/gauss_field.py:90-110, src/populate_field.py:11-29) turns a
density field into a Poisson catalogue of shape (N_p, 3), float32.  ``lognormal_mock``
runs the whole caller sequence of tests/create_lognormal.py:42-55 on the device
(Gaussian field -> irfftn -> exp(bias g) -> populate) without writing the spectrum,
optionally with a Zel'dovich redshift-space shift along z (``growth_rate``).

Differences from the reference, by design:
  * random numbers come from counter-based Philox4x32-10 streams of the integer seed
    (the reference uses NumPy's MT19937 and JAX's threefry), so a seed reproduces this
    generator's catalogue on any device and launch shape, not the reference's;
  * particles come out in C order of their cells (the reference sorts cells by decreasing
    density first; the catalogue is the same up to that permutation);
  * rates, normalisation and Poisson draws are FP64.
Each call allocates its outputs and scratch with torch; all arithmetic runs in
libpowspec_amd.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _device, _lib

__all__ = ["gaussian_field", "irfftn", "gaussian_mesh", "populate_field", "lognormal_mock"]


def _seed(seed) -> int:
    s = int(seed)
    if s < 0:
        raise ValueError("seed must be a non-negative integer")
    return s & 0xFFFFFFFFFFFFFFFF


def _pk_table(kf, Pkf, device):
    kf = np.asarray(kf.detach().cpu() if isinstance(kf, torch.Tensor) else kf, dtype=np.float64).ravel()
    pk = np.asarray(Pkf.detach().cpu() if isinstance(Pkf, torch.Tensor) else Pkf, dtype=np.float64).ravel()
    if kf.size < 2 or kf.size != pk.size:
        raise ValueError("kf and Pkf must be 1-D tables of the same length >= 2")
    if not np.all(np.diff(kf) > 0):
        raise ValueError("kf must be strictly increasing")
    return (torch.as_tensor(kf, device=device), torch.as_tensor(pk, device=device), int(kf.size))


def _field_ws(n, device):
    nbytes = C.c_size_t()
    _lib.check("pk_gaussian_workspace_size", _lib.load().pk_gaussian_workspace_size(int(n), C.byref(nbytes)))
    return _device.workspace(nbytes.value, device), nbytes.value


def gaussian_field(grid, kf, Pkf, Rayleigh_sampling, seed, BoxSize, *, device=None):
    """delta_k of src/gauss_field.py:5-80: complex64 (grid, grid, grid//2+1) on the GPU."""
    device = _device.require_gpu(device)
    n = int(grid)
    plan = _device.plan(n, device)
    lib = _lib.load()
    with torch.cuda.device(device):
        kf_t, pk_t, nk = _pk_table(kf, Pkf, device)
        ws, nbytes = _field_ws(n, device)
        spec = torch.empty((n // 2 + 1, n, n), dtype=torch.complex64, device=device)
        _lib.check("pk_gaussian_field", lib.pk_gaussian_field(
            plan.handle, float(BoxSize), _device.ptr(kf_t), _device.ptr(pk_t), nk, int(Rayleigh_sampling),
            _seed(seed), _device.ptr(spec), _device.ptr(ws), nbytes, _device.stream_ptr(device)))
    return spec.permute(1, 2, 0)


def irfftn(delta_k):
    """numpy/scipy irfftn of a half spectrum (n, n, n//2+1) -> float32 (n, n, n) mesh."""
    if not isinstance(delta_k, torch.Tensor):
        delta_k = torch.as_tensor(np.asarray(delta_k, dtype=np.complex64))
    device = _device.device_of(delta_k)
    if delta_k.dim() != 3 or delta_k.shape[0] != delta_k.shape[1] or delta_k.shape[2] != delta_k.shape[0] // 2 + 1:
        raise ValueError(f"expected a half spectrum of shape (n, n, n//2+1), got {tuple(delta_k.shape)}")
    n = int(delta_k.shape[0])
    plan = _device.plan(n, device)
    t = delta_k.detach()
    if t.device != device:
        t = t.to(device)
    if t.dtype != torch.complex64:
        t = t.to(torch.complex64)
    src = t.permute(2, 0, 1)              # device layout [kz][kx][ky]; a view for our own spectra
    if not src.is_contiguous():
        src = src.contiguous()
    with torch.cuda.device(device):
        scratch = torch.empty_like(src)
        mesh = torch.empty((n, n, n), dtype=torch.float32, device=device)
        _lib.check("pk_c2r", _lib.load().pk_c2r(plan.handle, _device.ptr(src), _device.ptr(scratch),
                                                _device.ptr(mesh), _device.stream_ptr(device)))
    return mesh


def gaussian_mesh(grid, kf, Pkf, Rayleigh_sampling, seed, BoxSize, *, component="delta", device=None):
    """irfftn(gaussian_field(...)) in one launch sequence (the spectrum is never stored);
    component='psi_z' gives the Zel'dovich displacement along z, irfftn(i k_z delta / k^2)."""
    device = _device.require_gpu(device)
    comp = {"delta": _lib.FIELD_DELTA, "psi_z": _lib.FIELD_PSI_Z}.get(component)
    if comp is None:
        raise ValueError(f"unknown component {component!r} (delta or psi_z)")
    n = int(grid)
    plan = _device.plan(n, device)
    with torch.cuda.device(device):
        kf_t, pk_t, nk = _pk_table(kf, Pkf, device)
        ws, nbytes = _field_ws(n, device)
        scratch = torch.empty((n // 2 + 1, n, n), dtype=torch.complex64, device=device)
        mesh = torch.empty((n, n, n), dtype=torch.float32, device=device)
        _run_mesh(plan, float(BoxSize), kf_t, pk_t, nk, int(Rayleigh_sampling), _seed(seed), comp, scratch, mesh,
                  ws, nbytes, device)
    return mesh


def _run_mesh(plan, box, kf_t, pk_t, nk, rayleigh, seed, comp, scratch, mesh, ws, nbytes, device):
    _lib.check("pk_gaussian_mesh", _lib.load().pk_gaussian_mesh(
        plan.handle, box, _device.ptr(kf_t), _device.ptr(pk_t), nk, rayleigh, seed, comp, _device.ptr(scratch),
        _device.ptr(mesh), _device.ptr(ws), nbytes, _device.stream_ptr(device)))


def populate_field(rho, n_bins, box_size, density, seed, *, bias=0.0, psi_z=None, growth_rate=0.0):
    """Poisson catalogue of a density field (src/gauss_field.py:90-110): float32 (N_p, 3).

    rates = rho (box_size/n_bins)^3 density / mean(rho); with ``bias`` != 0 the input is a
    Gaussian field g and rho = exp(bias g) (tests/create_lognormal.py:50).  ``psi_z`` (a
    mesh of z displacements) with ``growth_rate`` shifts particles along z by
    growth_rate * psi_z of their cell."""
    device = _device.device_of(rho, psi_z)
    mesh = _device.as_mesh(rho, device, "rho")
    n = int(mesh.shape[0])
    if int(n_bins) != n:
        raise ValueError(f"n_bins={n_bins} does not match the field's side {n}")
    psi = None
    if psi_z is not None and float(growth_rate) != 0.0:
        psi = _device.as_mesh(psi_z, device, "psi_z")
        if psi.shape[0] != n:
            raise ValueError("psi_z must have the field's shape")
    return _populate(mesh, n, float(box_size), float(density), float(bias), _seed(seed), psi, float(growth_rate),
                     device)


def _populate(mesh, n, box, density, bias, seed, psi, growth, device):
    lib = _lib.load()
    nbytes = C.c_size_t()
    _lib.check("pk_populate_workspace_size", lib.pk_populate_workspace_size(n, C.byref(nbytes)))
    with torch.cuda.device(device):
        ws = _device.workspace(nbytes.value, device)
        total = torch.zeros(1, dtype=torch.int64, device=device)
        stream = _device.stream_ptr(device)
        _lib.check("pk_populate_count", lib.pk_populate_count(
            _device.ptr(mesh), n, box, density, bias, seed, _device.ptr(ws), nbytes.value, _device.ptr(total), stream))
        npart = int(total.item())
        pos = torch.empty((npart, 3), dtype=torch.float32, device=device)
        _lib.check("pk_populate_place", lib.pk_populate_place(
            _device.ptr(mesh), _device.ptr(psi), growth, n, box, bias, seed, _device.ptr(ws), nbytes.value,
            _device.ptr(pos) if npart else None, npart, stream))
    return pos


def lognormal_mock(grid, kf, Pkf, box_size, density, bias, seed, *, Rayleigh_sampling=0, growth_rate=0.0,
                   return_fields=False, device=None):
    """The caller sequence of tests/create_lognormal.py:42-55 on the GPU: a Gaussian field
    of P(k) (seed), exp(bias g) normalised, Poisson population (seed + 1, as the caller's
    PRNGKey(seed + 1)), and with growth_rate != 0 a redshift-space shift
    z += growth_rate * psi_z.  Returns positions (N_p, 3) float32 in [0, box_size), and
    with return_fields=True also the Gaussian mesh (and psi_z mesh or None)."""
    if float(bias) == 0.0:
        raise ValueError("bias must be non-zero (rho = exp(bias * g))")
    device = _device.require_gpu(device)
    n = int(grid)
    plan = _device.plan(n, device)
    box = float(box_size)
    s = _seed(seed)
    with torch.cuda.device(device):
        kf_t, pk_t, nk = _pk_table(kf, Pkf, device)
        ws, nbytes = _field_ws(n, device)
        scratch = torch.empty((n // 2 + 1, n, n), dtype=torch.complex64, device=device)
        g = torch.empty((n, n, n), dtype=torch.float32, device=device)
        _run_mesh(plan, box, kf_t, pk_t, nk, int(Rayleigh_sampling), s, _lib.FIELD_DELTA, scratch, g, ws, nbytes,
                  device)
        psi = None
        if float(growth_rate) != 0.0:
            psi = torch.empty((n, n, n), dtype=torch.float32, device=device)
            _run_mesh(plan, box, kf_t, pk_t, nk, int(Rayleigh_sampling), s, _lib.FIELD_PSI_Z, scratch, psi, ws,
                      nbytes, device)
        del scratch
        pos = _populate(g, n, box, float(density), float(bias), (s + 1) & 0xFFFFFFFFFFFFFFFF, psi,
                        float(growth_rate), device)
    if return_fields:
        return pos, g, psi
    return pos
