"""k-shell binning tables: exact integer thresholds on m = kx^2 + ky^2 + kz^2.

The reference bins |k| = sqrt(m) with jnp.histogram (src/correlations.py:42-48):
bin index = searchsorted(edges, k, side='right'), k == edges[-1] goes to the last bin,
k < edges[0] and k > edges[-1] are dropped.  Doing that with a device sqrt would let
rounding move modes that sit exactly on an edge (0.2-0.8% of modes are at integer |k|).
Instead the edges are turned, on the host in FP64, into integer thresholds

    thr[i] = min{ m : sqrt(m) >= e_i }          (i < nb)
    thr[nb] = min{ m : sqrt(m) >  e_nb }         (the closed last edge)

so that the GPU only compares integers:  bin(m) = #{i < nb : thr[i] <= m} - 1, valid when
thr[0] <= m < thr[nb].  sqrt here is numpy's FP64 sqrt (correctly rounded, monotone), the
same comparison the FP64 histogram makes, so bin indices and mode counts are bit-exact.
jstart[j] = #{i < nb : thr[i] <= j^2} - 1 lets the kernel start the search at isqrt(m).
"""
from __future__ import annotations

import math
from collections import OrderedDict

import numpy as np


def max_m(n: int) -> int:
    """Largest m = kx^2+ky^2+kz^2 on the stored half spectrum (|k_a| <= n/2)."""
    h = n // 2
    return 3 * h * h


def jstart_len(n: int) -> int:
    return math.isqrt(max_m(n)) + 2


def _min_m_ge(e: np.ndarray, never: int, strict: bool) -> np.ndarray:
    """Vectorised min{m in [0, never) : sqrt(m) >= e} (> e if strict); `never` if none."""
    e = np.asarray(e, np.float64)
    out = np.full(e.shape, never, np.int64)
    fin = np.isfinite(e)
    neg = fin & ((e < 0) | ((e == 0) & (not strict)))
    out[neg] = 0
    todo = fin & ~neg
    if np.any(todo):
        ev = e[todo]
        guess = np.clip(np.ceil(ev * ev), 0, never).astype(np.int64)
        cmp = (lambda s, x: s > x) if strict else (lambda s, x: s >= x)
        for _ in range(8):   # fix-ups around the float guess
            m_lo = np.maximum(guess - 1, 0)
            down = (guess > 0) & cmp(np.sqrt(m_lo.astype(np.float64)), ev)
            up = (guess < never) & ~cmp(np.sqrt(guess.astype(np.float64)), ev)
            if not (down.any() or up.any()):
                break
            guess = np.where(down, guess - 1, np.where(up, guess + 1, guess))
        guess = np.clip(guess, 0, never)
        out[todo] = guess
    # NaN edges never match (numpy searchsorted orders NaN last); +inf likewise
    return out


def thresholds_from_grid_edges(edges: np.ndarray, n: int):
    """Integer tables for edges given in units of k_F (kedges = k_edges / kF).

    Returns (thr int32[nb+1], jstart int32[jstart_len(n)]).  Raises ValueError for fewer
    than two edges or decreasing edges.
    """
    e = np.asarray(edges, np.float64).reshape(-1)
    if e.size < 2:
        raise ValueError("k_edges needs at least two edges")
    with np.errstate(invalid="ignore"):
        d = np.diff(e)
    if np.any(d[np.isfinite(d)] < 0):
        raise ValueError("k_edges must be monotonically non-decreasing")
    nb = e.size - 1
    never = max_m(n) + 1
    thr = np.empty(nb + 1, np.int64)
    thr[:nb] = _min_m_ge(e[:nb], never, strict=False)
    thr[nb] = _min_m_ge(e[nb:], never, strict=True)[0]
    thr = np.minimum(thr, never)
    js = np.arange(jstart_len(n), dtype=np.int64)
    jstart = np.searchsorted(thr[:nb], js * js, side="right") - 1
    return thr.astype(np.int32), jstart.astype(np.int32)


def slots_for_m(m: np.ndarray, thr: np.ndarray) -> np.ndarray:
    """Host evaluation of the device rule (for tests): slot of each m, -1 if dropped."""
    nb = thr.size - 1
    m = np.asarray(m, np.int64)
    idx = np.searchsorted(thr[:nb].astype(np.int64), m, side="right") - 1
    ok = (idx >= 0) & (m < thr[nb])
    return np.where(ok, idx, -1)


_MULT: dict = {}


def mode_multiplicity(n: int) -> np.ndarray:
    """mult[m] = number of stored half-spectrum entries (kx, ky in [0, n), kz in [0, n/2],
    signed wavenumbers with Nyquist +n/2) with kx^2 + ky^2 + kz^2 = m; cached per n."""
    hit = _MULT.get(n)
    if hit is not None:
        return hit
    ki = np.arange(n)
    ki = np.where(ki > n // 2, ki - n, ki).astype(np.int64)
    hist_xy = np.bincount((ki[:, None] ** 2 + ki[None, :] ** 2).reshape(-1))
    mult = np.zeros(max_m(n) + 1, np.int64)
    for kz in range(n // 2 + 1):
        mult[kz * kz:kz * kz + hist_xy.size] += hist_xy
    _MULT[n] = mult
    return mult


_MULT_FULL: dict = {}


def full_grid_multiplicity(n: int) -> np.ndarray:
    """mult[m] = number of cells of the full n^3 separation grid (signed coordinates, +n/2
    kept) with sx^2 + sy^2 + sz^2 = m: the histogram population of the configuration-space
    estimators (src/correlations.py:157-177); cached per n."""
    hit = _MULT_FULL.get(n)
    if hit is not None:
        return hit
    ki = np.arange(n)
    ki = np.where(ki > n // 2, ki - n, ki).astype(np.int64)
    hist_xy = np.bincount((ki[:, None] ** 2 + ki[None, :] ** 2).reshape(-1))
    zc = np.bincount(ki * ki)                   # multiplicity of each sz^2
    mult = np.zeros(max_m(n) + 1, np.int64)
    for v in np.nonzero(zc)[0]:
        mult[v:v + hist_xy.size] += zc[v] * hist_xy
    _MULT_FULL[n] = mult
    return mult


def bin_modes(thr: np.ndarray, n: int, full: bool = False):
    """Exact per-bin mode counts (int64) and sums of |k| in units of k_F (FP64) over the
    stored half spectrum, from the integer thresholds: every entry counts once, the k = 0
    entry included when its bin exists.  full=True: over the full n^3 grid instead."""
    mult = full_grid_multiplicity(n) if full else mode_multiplicity(n)
    nb = thr.size - 1
    m = np.arange(mult.size, dtype=np.int64)
    slot = slots_for_m(m, thr)
    sel = slot >= 0
    counts = np.bincount(slot[sel], weights=mult[sel], minlength=nb)[:nb]
    ksum = np.bincount(slot[sel], weights=mult[sel] * np.sqrt(m[sel].astype(np.float64)), minlength=nb)[:nb]
    return np.rint(counts).astype(np.int64), ksum


def fundamental_grid_edges(n: int) -> np.ndarray:
    """Edges 1, 2, ..., kmax+1 (k_F units): bin j = [j, j+1) for j >= 1, the slots of
    src/correlations.py:95-109 after the DC slot is dropped; kmax = int32(sqrt(3 mid^2))."""
    kmax = math.isqrt(max_m(n))
    return np.arange(1, kmax + 2, dtype=np.float64)


def grid_edges(k_edges, box_size: float) -> np.ndarray:
    """kedges = k_edges / kF with kF = 2 pi / box_size, FP64 (src/correlations.py:12, :42)."""
    if hasattr(k_edges, "detach"):
        k_edges = k_edges.detach().cpu().double().numpy()
    kF = 2.0 * np.pi / float(box_size)
    return np.asarray(k_edges, np.float64).reshape(-1) / kF


class BinTable:
    """Device-resident bin tables of one (n, edges): thr int32[nb+1], jst int32, exact mode
    counts int64[nb] and sums of |k| (k_F units) FP64[nb]."""

    def __init__(self, thr, jst, nb, modes, ksum):
        self.thr, self.jst, self.nb, self.modes, self.ksum = thr, jst, nb, modes, ksum


class TableCache:
    """Bin tables keyed by (device, n, edges).  Edge arrays that give the same integer
    thresholds (e.g. the same k_F multiples from different box sizes) share one device
    table; both levels are LRUs."""

    def __init__(self, capacity: int = 4096, tables: int = 256, full: bool = False):
        self.full = full
        self.capacity = capacity
        self.tables = tables
        self._d: OrderedDict = OrderedDict()
        self._t: OrderedDict = OrderedDict()

    def get(self, device, n: int, edges: np.ndarray):
        import torch
        key = (str(device), int(n), edges.tobytes())
        hit = self._d.get(key)
        if hit is not None:
            self._d.move_to_end(key)
            return hit
        thr, jst = thresholds_from_grid_edges(edges, n)
        tkey = (str(device), int(n), thr.tobytes())
        val = self._t.get(tkey)
        if val is None:
            modes, ksum = bin_modes(thr, n, self.full)
            val = BinTable(torch.from_numpy(thr).to(device), torch.from_numpy(jst).to(device), thr.size - 1,
                           torch.from_numpy(modes).to(device), torch.from_numpy(ksum).to(device))
            self._t[tkey] = val
            if len(self._t) > self.tables:
                self._t.popitem(last=False)
        else:
            self._t.move_to_end(tkey)
        self._d[key] = val
        if len(self._d) > self.capacity:
            self._d.popitem(last=False)
        return val


TABLES = TableCache()
# tables over the full separation grid (configuration-space multipoles)
TABLES_FULL = TableCache(full=True)
