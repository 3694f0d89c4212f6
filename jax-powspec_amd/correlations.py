"""Power-spectrum multipoles -- drop-in for the reference's src/correlations.py P_l path.

``powspec_vec`` (src/correlations.py:7-56) and ``powspec_vec_fundamental`` (:60-117) keep
the reference signatures and return conventions: (k3D, Pk3D[nb, 3] float32, Nmodes3D),
P_l for l = 0, 2, 4 with the line of sight along z, CIC window deconvolution, every stored
half-spectrum entry counted once, NaN in empty bins.  Extensions are keyword-only:
``mas`` (window exponent: 'CIC' p=2 as the reference, 'TSC' p=3, 'NGP' p=1, 'NONE' p=0),
``shot_noise`` (subtracted from P_0, as callers do at tests/correlations.py:83) and
``full=True`` (FP64 sums, int64 counts, mode-averaged k).

``estimate`` runs the whole path from particles (assignment -> contrast -> P_l) in one
call with the contrast folded into the FFT's first read.

Differences from the reference, by design:
  * bins are decided on exact integer thresholds derived from FP64 edges (binning.py),
  * per-bin sums are FP64 on the device, outputs are float32 like the reference's Pk3D,
  * powspec_vec_fundamental returns the mode-averaged k (the reference's k at :100 uses
    ``.set``, i.e. whichever mode wrote last),
  * decreasing k_edges raise ValueError (jnp.histogram silently misbins them).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _device, _lib, binning

__all__ = ["powspec_vec", "powspec_vec_fundamental", "xi_vec", "xi_vec_fundamental", "xi_vec_coords",
           "s_edges_conv", "bispec", "estimate", "rfftn", "Estimator", "WINDOW_P"]

WINDOW_P = {"NONE": 0, "NGP": 1, "CIC": 2, "CIC_REF": 2, "TSC": 3}


def _window_p(mas) -> int:
    if isinstance(mas, (int, np.integer)):
        p = int(mas)
    else:
        key = "NONE" if mas is None else str(mas).upper()
        if key not in WINDOW_P:
            raise ValueError(f"unknown window {mas!r}")
        p = WINDOW_P[key]
    if not 0 <= p <= 3:
        raise ValueError(f"window exponent {p} not in 0..3")
    return p


class _Outputs:
    def __init__(self, nb: int, device, want_k: bool):
        self.sums = torch.empty((nb, 3), dtype=torch.float64, device=device)
        self.counts = torch.empty(nb, dtype=torch.int64, device=device)
        self.pk = torch.empty((nb, 3), dtype=torch.float32, device=device)
        self.nmodes = torch.empty(nb, dtype=torch.float32, device=device)
        self.kmean = torch.empty(nb, dtype=torch.float64, device=device) if want_k else None


def _powspec(delta, box_size, edges_grid, *, p, shot_noise, flags, stats=None):
    device = _device.device_of(delta)
    mesh = _device.as_mesh(delta, device)
    n = mesh.shape[0]
    plan = _device.plan(n, device)
    tab = binning.TABLES.get(device, n, edges_grid)
    nb = tab.nb
    lib = _lib.load()
    out = _Outputs(nb, device, bool(flags & _lib.FLAG_KMEAN))
    nbytes = C.c_size_t()
    with torch.cuda.device(device):
        _lib.check("pk_powspec_workspace_size", lib.pk_powspec_workspace_size(plan.handle, nb, C.byref(nbytes)))
        ws = _device.workspace(nbytes.value, device)
        spec = torch.empty((n // 2 + 1, n, n, 2), dtype=torch.float32, device=device)
        _lib.check("pk_powspec_mesh", lib.pk_powspec_mesh(
            plan.handle, _device.ptr(mesh), _device.ptr(stats), _device.ptr(tab.thr), _device.ptr(tab.jst),
            _device.ptr(tab.modes), _device.ptr(tab.ksum), nb, p, flags,
            float(box_size), float(shot_noise), _device.ptr(spec), _device.ptr(out.sums), _device.ptr(out.counts),
            _device.ptr(out.pk), _device.ptr(out.nmodes), _device.ptr(out.kmean), _device.ptr(ws), nbytes.value,
            _device.stream_ptr(device)))
    return out, device


def _extras(out, kF):
    d = dict(sums=out.sums, counts=out.counts)
    if out.kmean is not None:
        d["kmean"] = out.kmean * kF
    return d


def powspec_vec(delta, box_size, k_edges, *, mas="CIC", shot_noise=0.0, full=False):
    """P_l(k), l = 0, 2, 4, of a density-contrast mesh in user k bins (src/correlations.py:7-56).

    Returns (k3D, Pk3D, Nmodes3D): bin centres 0.5 (e_i + e_{i+1}) (:54), float32 (nb, 3)
    multipoles normalised by (box/n^2)^3 (x5, x9 for l = 2, 4; :50-52) and float32 mode
    counts.  With full=True a 4th element holds FP64 sums, int64 counts and the
    mode-averaged k.
    """
    e = binning.grid_edges(k_edges, box_size)
    flags = _lib.FLAG_KMEAN if full else 0
    out, device = _powspec(delta, box_size, e, p=_window_p(mas), shot_noise=shot_noise, flags=flags)
    kF = 2.0 * np.pi / float(box_size)
    k3D = torch.as_tensor((0.5 * (e[1:] + e[:-1]) * kF).astype(np.float32), device=device)
    res = (k3D, out.pk, out.nmodes)
    return res + (_extras(out, kF),) if full else res


def powspec_vec_fundamental(delta, box_size, *, mas="CIC", shot_noise=0.0, full=False):
    """P_l(k) in k_F-wide bins [j, j+1) k_F, j = 1..kmax (src/correlations.py:60-117).

    kmax = int32(sqrt(3 (n/2)^2)) (:69); the DC slot is dropped (:106-109).  k3D is the
    mode-averaged k of each bin.
    """
    device = _device.device_of(delta)
    n = int(delta.shape[0])
    e = binning.fundamental_grid_edges(n)
    flags = _lib.FLAG_KMEAN
    out, device = _powspec(delta, box_size, e, p=_window_p(mas), shot_noise=shot_noise, flags=flags)
    kF = 2.0 * np.pi / float(box_size)
    k3D = (out.kmean * kF).to(torch.float32)
    res = (k3D, out.pk, out.nmodes)
    return res + (_extras(out, kF),) if full else res


def _xi(delta, edges_grid, p, flags):
    device = _device.device_of(delta)
    mesh = _device.as_mesh(delta, device)
    n = mesh.shape[0]
    plan = _device.plan(n, device)
    tab = binning.TABLES_FULL.get(device, n, edges_grid)
    nb = tab.nb
    if nb > _lib.XI_MAX_BINS:
        raise ValueError(f"xi supports at most {_lib.XI_MAX_BINS} bins (got {nb})")
    lib = _lib.load()
    nbytes = C.c_size_t()
    with torch.cuda.device(device):
        _lib.check("pk_xi_workspace_size", lib.pk_xi_workspace_size(plan.handle, nb, C.byref(nbytes)))
        ws = _device.workspace(nbytes.value, device)
        sums = torch.empty((nb, 3), dtype=torch.float64, device=device)
        xi = torch.empty((nb, 3), dtype=torch.float32, device=device)
        nm = torch.empty(nb, dtype=torch.float32, device=device)
        _lib.check("pk_xi_mesh", lib.pk_xi_mesh(
            plan.handle, _device.ptr(mesh), _device.ptr(tab.thr), _device.ptr(tab.jst), _device.ptr(tab.modes), nb,
            p, flags, _device.ptr(sums), _device.ptr(xi), _device.ptr(nm), _device.ptr(ws), nbytes.value,
            _device.stream_ptr(device)))
    return xi, nm, sums, tab, device


def s_edges_conv(dims, box_size, s_edges):
    """k-space form of separation edges, kF s n / L (src/correlations.py:269-272)."""
    kF = 2.0 * np.pi / float(box_size)
    return kF * np.asarray(s_edges, np.float64) * int(dims) / float(box_size)


def xi_vec_coords(dims, box_size, k_edges):
    """Bin centres of xi_vec for k-space edges (src/correlations.py:262-266)."""
    kF = 2.0 * np.pi / float(box_size)
    ke = np.asarray(k_edges, np.float64) / kF
    return 0.5 * (ke[1:] + ke[:-1]) * (float(box_size) / int(dims))


def xi_vec(delta, box_size, s_edges, *, mas="CIC", full=False):
    """Correlation-function multipoles xi_l(s), l = 0, 2, 4 (src/correlations.py:119-187).

    xi(s) = irfftn(|rfftn(delta) W|^2) over the full n^3 separation grid, binned by |s|
    with the reference's histogram rules (edges s_edges n / L in grid units), mu = s_z/|s|.
    Returns (r3D, xi3D[nb, 3] float32, Nmodes3D float32 with inf in empty bins, where xi is
    0).  The bin holding s = 0 (an edge at or below 0) has NaN l = 2, 4, as the reference.
    """
    if hasattr(s_edges, "detach"):
        s_edges = s_edges.detach().cpu().double().numpy()
    n = int(delta.shape[0])
    e = np.asarray(s_edges, np.float64).reshape(-1) * n / float(box_size)
    xi, nm, sums, tab, device = _xi(delta, e, _window_p(mas), 0)
    r = torch.as_tensor((0.5 * (e[1:] + e[:-1]) * (float(box_size) / n)).astype(np.float32), device=device)
    res = (r, xi, nm)
    return res + ({"sums": sums, "counts": tab.modes},) if full else res


def xi_vec_fundamental(delta, box_size, *, mas="CIC", full=False):
    """xi_l(s) in grid-spacing-wide shells [j, j+1), j = 1..kmax (src/correlations.py:191-260):
    r3D is the cell-averaged |s| of each shell; empty shells give NaN."""
    n = int(delta.shape[0])
    e = binning.fundamental_grid_edges(n)
    xi, nm, sums, tab, device = _xi(delta, e, _window_p(mas), _lib.XI_EMPTY_NAN)
    with np.errstate(invalid="ignore", divide="ignore"):
        modes = tab.modes.cpu().numpy().astype(np.float64)
        r = tab.ksum.cpu().numpy() / modes * (float(box_size) / n)
    res = (torch.as_tensor(r.astype(np.float32), device=device), xi, nm)
    return res + ({"sums": sums, "counts": tab.modes},) if full else res


def bispec_shells(n, box_size, k1, k2, theta):
    """k_all = (k1, k2, k3(theta)) (src/correlations.py:352-356) and the integer shells
    lo <= m < hi equivalent to k_i/kF - 1 <= |k| < k_i/kF + 1 (:358-359, :377) in FP64."""
    kF = 2.0 * np.pi / float(box_size)
    theta = np.asarray(theta.detach().cpu() if hasattr(theta, "detach") else theta, np.float64).reshape(-1)
    k3 = np.sqrt((k2 * np.sin(theta)) ** 2 + (k2 * np.cos(theta) + k1) ** 2)
    k_all = np.concatenate([[float(k1), float(k2)], k3])
    never = binning.max_m(n) + 1
    lo = binning._min_m_ge((k_all - kF) / kF, never, strict=False)
    hi = binning._min_m_ge((k_all + kF) / kF, never, strict=False)
    return k_all, theta, lo.astype(np.int32), hi.astype(np.int32)


# sum I_i^2 and sum I1 I2 I_i of the bispectrum's shell indicators per (device, n, shells):
# they do not depend on the field, so later calls skip the indicator transforms
_BISPEC_I_SUMS: dict = {}


def bispec(delta, box_size, k1, k2, theta, *, mas="CIC", full=False, reuse_shell_sums=True):
    """Bispectrum and reduced bispectrum of (k1, k2, theta) triangles (src/correlations.py:
    334-458).  Returns (k_all[bins+2], Pk[bins+2], theta, B[bins], Q[bins]) as float32
    device tensors (k_all and theta from the host computation).  mas: 'CIC' (the
    reference's (x / sin x)^2 per axis on delta_k) or 'NONE'.  The shell indicators' sums
    are kept per (device, n, shells) after the first call and reused (reuse_shell_sums)."""
    key = "NONE" if mas is None else str(mas).upper()
    if key not in ("CIC", "CIC_REF", "NONE"):
        raise ValueError(f"bispec supports mas='CIC' or 'NONE', got {mas!r}")
    p = 0 if key == "NONE" else 1
    device = _device.device_of(delta)
    mesh = _device.as_mesh(delta, device)
    n = mesh.shape[0]
    plan = _device.plan(n, device)
    k_all, th, lo, hi = bispec_shells(n, box_size, k1, k2, theta)
    ns = k_all.size
    if ns > _lib.BISPEC_MAX_SHELLS:
        raise ValueError(f"at most {_lib.BISPEC_MAX_SHELLS - 2} theta bins")
    lib = _lib.load()
    nbytes = C.c_size_t()
    key = (str(device), n, lo.tobytes(), hi.tobytes())
    isums = _BISPEC_I_SUMS.get(key) if reuse_shell_sums else None
    with torch.cuda.device(device):
        _lib.check("pk_bispec_workspace_size", lib.pk_bispec_workspace_size(plan.handle, ns, C.byref(nbytes)))
        ws = _device.workspace(nbytes.value, device)
        sums = torch.empty((ns, 4), dtype=torch.float64, device=device)
        pk = torch.empty(ns, dtype=torch.float32, device=device)
        bk = torch.empty(max(ns - 2, 1), dtype=torch.float32, device=device)
        qk = torch.empty(max(ns - 2, 1), dtype=torch.float32, device=device)
        lo_h = (C.c_int32 * ns)(*lo.tolist())
        hi_h = (C.c_int32 * ns)(*hi.tolist())
        _lib.check("pk_bispec_mesh", lib.pk_bispec_mesh(
            plan.handle, _device.ptr(mesh), C.cast(lo_h, C.c_void_p), C.cast(hi_h, C.c_void_p), ns, p,
            float(box_size), _device.ptr(isums), _device.ptr(sums), _device.ptr(pk), _device.ptr(bk), _device.ptr(qk),
            _device.ptr(ws), nbytes.value, _device.stream_ptr(device)))
        if reuse_shell_sums and isums is None:
            if len(_BISPEC_I_SUMS) > 256:
                _BISPEC_I_SUMS.clear()
            _BISPEC_I_SUMS[key] = sums[:, [1, 3]].contiguous()
    f32 = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=device)
    res = (f32(k_all), pk, f32(th), bk[:ns - 2], qk[:ns - 2])
    return res + ({"sums": sums, "shell_lo": lo, "shell_hi": hi},) if full else res


def rfftn(delta, *, contrast_stats=None):
    """Unnormalised forward R2C FFT of a float32 mesh (src/correlations.py:22).

    Returns a complex64 tensor of shape (n, n, n/2+1) equal to numpy.fft.rfftn(delta)
    (a permuted view of the library's [kz][kx][ky] layout).
    """
    device = _device.device_of(delta)
    mesh = _device.as_mesh(delta, device)
    n = mesh.shape[0]
    plan = _device.plan(n, device)
    spec = torch.empty((n // 2 + 1, n, n, 2), dtype=torch.float32, device=device)
    with torch.cuda.device(device):
        _lib.call("pk_r2c", plan.handle, _device.ptr(mesh), _device.ptr(contrast_stats), _device.ptr(spec),
                  _device.stream_ptr(device))
    return torch.view_as_complex(spec).permute(1, 2, 0)


def estimate(x, y, z, w=None, *, box_size, n_bins, k_edges=None, mas="CIC_REF", wrap=True, window=None,
             xmin=0.0, ymin=0.0, zmin=0.0, shot_noise=0.0, tile=None, full=False, stats=None):
    """Whole path from particles: assignment -> delta/mean - 1 -> P_l(k) (l = 0, 2, 4).

    Equivalent to cic_mas_vec (mas='CIC_REF') + ``delta /= delta.mean(); delta -= 1``
    (tests/lognormal.py:79-80) + powspec_vec (or powspec_vec_fundamental when k_edges is
    None) + ``pk[:, 0] - shot_noise``, with the contrast folded into the FFT's first read
    and the k = 0 mode taken as exactly 0.  window: deconvolution exponent (default: 2 for
    CIC variants, 3 for TSC).  Returns (k3D, Pk3D, Nmodes3D[, extras]).
    """
    mas = mas.upper()
    if mas not in _lib.STENCIL:
        raise ValueError(f"unknown mass-assignment scheme {mas!r}")
    n = int(n_bins)
    device = _device.device_of(x, y, z, w)
    plan = _device.plan(n, device)
    xs, sx = _device.as_f32_vector(x, device, "x")
    ys, sy = _device.as_f32_vector(y, device, "y")
    zs, sz = _device.as_f32_vector(z, device, "z")
    np_ = xs.numel()
    if ys.numel() != np_ or zs.numel() != np_:
        raise ValueError("x, y, z must have the same length")
    ws_, sw = (None, 1)
    if w is not None:
        ws_, sw = _device.as_f32_vector(w, device, "w")
    fundamental = k_edges is None
    e = binning.fundamental_grid_edges(n) if fundamental else binning.grid_edges(k_edges, box_size)
    tab = binning.TABLES.get(device, n, e)
    nb = tab.nb
    p = _window_p(window if window is not None else ("TSC" if mas == "TSC" else "CIC"))
    flags = _lib.FLAG_KMEAN if (full or fundamental) else 0
    out = _Outputs(nb, device, bool(flags & _lib.FLAG_KMEAN))
    tptr, _keep = _lib.tile_arg(tile)
    lib = _lib.load()
    nbytes = C.c_size_t()
    stencil = _lib.STENCIL[mas]
    with torch.cuda.device(device):
        _lib.check("pk_estimate_workspace_size",
                   lib.pk_estimate_workspace_size(plan.handle, np_, stencil, tptr, nb, C.byref(nbytes)))
        ws = _device.workspace(nbytes.value, device)
        _lib.check("pk_estimate", lib.pk_estimate(
            plan.handle, _device.ptr(xs), sx, _device.ptr(ys), sy, _device.ptr(zs), sz, _device.ptr(ws_), sw, np_,
            _device.as_f32_scalar(xmin), _device.as_f32_scalar(ymin), _device.as_f32_scalar(zmin), float(box_size),
            stencil, _lib.WRAP_VEC_TRUE if wrap else _lib.WRAP_VEC_FALSE, tptr, _device.ptr(tab.thr),
            _device.ptr(tab.jst), _device.ptr(tab.modes), _device.ptr(tab.ksum), nb, p, flags, float(shot_noise), _device.ptr(out.sums), _device.ptr(out.counts),
            _device.ptr(out.pk), _device.ptr(out.nmodes), _device.ptr(out.kmean), _device.ptr(stats),
            _device.ptr(ws), nbytes.value, _device.stream_ptr(device)))
    kF = 2.0 * np.pi / float(box_size)
    if fundamental:
        k3D = (out.kmean * kF).to(torch.float32)
    else:
        k3D = torch.as_tensor((0.5 * (e[1:] + e[:-1]) * kF).astype(np.float32), device=device)
    res = (k3D, out.pk, out.nmodes)
    return res + (_extras(out, kF),) if full else res


class Estimator:
    """Preallocated particles -> P_l path for repeated calls with the same (n, edges).

    Holds the plan, the device bin tables, the outputs and a workspace sized for up to
    ``max_particles``; ``run`` is a single pk_estimate launch sequence on the current
    stream with no allocation.  Outputs are overwritten by every call.
    """

    def __init__(self, *, box_size, n_bins, max_particles, k_edges=None, mas="CIC_REF", wrap=True, window=None,
                 tile=None, shot_noise=0.0, device=None, want_k=False):
        self.device = _device.require_gpu(device)
        self.n = int(n_bins)
        self.box = float(box_size)
        self.mas = mas.upper()
        if self.mas not in _lib.STENCIL:
            raise ValueError(f"unknown mass-assignment scheme {mas!r}")
        self.stencil = _lib.STENCIL[self.mas]
        self.wrap_mode = _lib.WRAP_VEC_TRUE if wrap else _lib.WRAP_VEC_FALSE
        self.plan = _device.plan(self.n, self.device)
        self.fundamental = k_edges is None
        self.edges = (binning.fundamental_grid_edges(self.n) if self.fundamental
                      else binning.grid_edges(k_edges, self.box))
        self.tables = binning.TABLES.get(self.device, self.n, self.edges)
        self.nb = self.tables.nb
        self.p = _window_p(window if window is not None else ("TSC" if self.mas == "TSC" else "CIC"))
        self.flags = _lib.FLAG_KMEAN if (want_k or self.fundamental) else 0
        self.out = _Outputs(self.nb, self.device, bool(self.flags & _lib.FLAG_KMEAN))
        self.stats = torch.zeros(_lib.STATS_LEN, dtype=torch.float64, device=self.device)
        self.shot_noise = float(shot_noise)
        self.max_particles = int(max_particles)
        self._tile_ptr, self._tile_keep = _lib.tile_arg(tile)
        lib = _lib.load()
        nbytes = C.c_size_t()
        with torch.cuda.device(self.device):
            _lib.check("pk_estimate_workspace_size", lib.pk_estimate_workspace_size(
                self.plan.handle, self.max_particles, self.stencil, self._tile_ptr, self.nb, C.byref(nbytes)))
        self.ws_bytes = nbytes.value
        self.ws = _device.workspace(self.ws_bytes, self.device)
        kF = 2.0 * np.pi / self.box
        self.kF = kF
        self.k_centres = torch.as_tensor((0.5 * (self.edges[1:] + self.edges[:-1]) * kF).astype(np.float32),
                                         device=self.device)

    def run(self, x, y, z, w=None, *, stream=None):
        xs, sx = _device.as_f32_vector(x, self.device, "x")
        ys, sy = _device.as_f32_vector(y, self.device, "y")
        zs, sz = _device.as_f32_vector(z, self.device, "z")
        np_ = xs.numel()
        if np_ > self.max_particles:
            raise ValueError(f"{np_} particles exceed the preallocated {self.max_particles}")
        ws_, sw = (None, 1)
        if w is not None:
            ws_, sw = _device.as_f32_vector(w, self.device, "w")
        o = self.out
        _lib.check("pk_estimate", _lib.load().pk_estimate(
            self.plan.handle, _device.ptr(xs), sx, _device.ptr(ys), sy, _device.ptr(zs), sz, _device.ptr(ws_), sw,
            np_, 0.0, 0.0, 0.0, self.box, self.stencil, self.wrap_mode, self._tile_ptr, _device.ptr(self.tables.thr),
            _device.ptr(self.tables.jst), _device.ptr(self.tables.modes), _device.ptr(self.tables.ksum), self.nb, self.p, self.flags, self.shot_noise, _device.ptr(o.sums),
            _device.ptr(o.counts), _device.ptr(o.pk), _device.ptr(o.nmodes), _device.ptr(o.kmean),
            _device.ptr(self.stats), _device.ptr(self.ws), self.ws_bytes, _device.stream_ptr(self.device, stream)))
        return o.pk

    def k(self):
        if self.fundamental:
            return (self.out.kmean * self.kF).to(torch.float32)
        return self.k_centres
