"""CPU baseline: FP32 NumPy/SciPy restatement of the reference algorithm -- TEST/BENCH ONLY.

Used only by bench.py's ``cpu_baseline`` leg (and tests) as the CPU reference point; never
by the product path.  It follows the reference step by step in float32, as the reference
ran under JAX's default x32 mode:

  scatter   8 scatter-adds of the reference stencil as np.bincount   src/mas.py:142-151
  contrast  delta / mean - 1                                          tests/lognormal.py:79-80
  FFT       scipy.fft.rfftn(workers=ncores), complex64                src/correlations.py:22
  window    prod (x / sin x)^2 per axis on delta_k                    src/correlations.py:32,39
  binning   searchsorted + 4 weighted histograms (np.bincount)        src/correlations.py:45-48
  normalise S / N * (L / n^2)^3 * (1, 5, 9)                           src/correlations.py:50-52
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft

F32 = np.float32


def ncores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cic_mas_vec_f32(x, y, z, w, box_size, n):
    inv = F32(1.0) / (F32(box_size) / F32(n))
    pos = [(np.asarray(c, F32) - F32(0.0)) * inv for c in (x, y, z)]
    i = [p.astype(np.int32) for p in pos]
    d = [p - q.astype(F32) for p, q in zip(pos, i)]
    m = [F32(1.0) - v for v in d]
    ip = [(q + 1 + n) % n for q in i]
    w = np.asarray(w, F32)
    flat = np.zeros(n ** 3, np.float64)
    spec = [((0, 0, 0), m[0] * m[1] * m[2]), ((1, 0, 0), d[0] * m[1] * m[2]), ((0, 1, 0), m[0] * d[1] * m[2]),
            ((0, 0, 1), m[0] * m[1] * d[2]), ((1, 1, 0), d[0] * d[1] * m[2]), ((1, 0, 1), d[0] * m[1] * d[2]),
            ((0, 1, 1), m[0] * m[1] * d[2]), ((1, 1, 1), d[0] * d[1] * d[2])]
    for (sx, sy, sz), wt in spec:
        ix = ip[0] if sx else i[0]
        iy = ip[1] if sy else i[1]
        iz = ip[2] if sz else i[2]
        ok = (ix >= 0) & (ix < n) & (iy >= 0) & (iy < n) & (iz >= 0) & (iz < n)
        lin = (ix.astype(np.int64) * n + iy) * n + iz
        flat += np.bincount(lin[ok], weights=(wt * w)[ok], minlength=n ** 3)
    return flat.astype(F32).reshape(n, n, n)


def powspec_vec_f32(delta, box_size, k_edges, workers=None):
    delta = np.asarray(delta, F32)
    n = delta.shape[0]
    mid = n // 2
    kF = F32(2.0 * np.pi / box_size)
    dk = scipy.fft.rfftn(delta, workers=workers or ncores())
    r = np.arange(n)
    ki = np.where(r > mid, r - n, r).astype(F32)
    corr1 = (F32(1.0) / np.sinc(ki / F32(n)).astype(F32)) ** 2
    corr = corr1[:, None, None] * corr1[None, :, None] * corr1[None, None, : mid + 1]
    dk *= corr
    d2 = (dk.real * dk.real + dk.imag * dk.imag).astype(F32).reshape(-1)
    kx, ky, kz = ki[:, None, None], ki[None, :, None], ki[None, None, : mid + 1]
    k = np.sqrt(kx * kx + ky * ky + kz * kz).reshape(-1)
    with np.errstate(invalid="ignore", divide="ignore"):
        mu2 = np.where(k == 0, F32(0.0), (np.broadcast_to(kz, (n, n, mid + 1)).reshape(-1) / k) ** 2)
    e = (np.asarray(k_edges, F32) / kF).astype(F32)
    nb = e.size - 1
    idx = np.searchsorted(e, k, side="right")
    idx = np.where(k == e[-1], nb, idx)
    ok = (idx >= 1) & (idx <= nb)
    s = idx[ok] - 1
    P = d2[ok]
    m2 = mu2[ok]
    S0 = np.bincount(s, weights=P, minlength=nb)
    S2 = np.bincount(s, weights=P * (3 * m2 - 1) / 2, minlength=nb)
    S4 = np.bincount(s, weights=P * (35 * m2 * m2 - 30 * m2 + 3) / 8, minlength=nb)
    N = np.bincount(s, minlength=nb).astype(F32)
    norm = F32((box_size / n ** 2) ** 3)
    with np.errstate(invalid="ignore", divide="ignore"):
        Pk = np.stack([S0 / N * norm, S2 / N * 5 * norm, S4 / N * 9 * norm], 1).astype(F32)
    return 0.5 * (e[1:] + e[:-1]) * kF, Pk, N


def estimate_f32(x, y, z, box_size, n, k_edges, shot_noise=0.0, workers=None):
    mesh = cic_mas_vec_f32(x, y, z, np.ones(np.asarray(x).shape, F32), box_size, n)
    mesh /= mesh.mean()
    mesh -= F32(1.0)
    k, Pk, N = powspec_vec_f32(mesh, box_size, k_edges, workers)
    Pk[:, 0] -= F32(shot_noise)
    return k, Pk, N
