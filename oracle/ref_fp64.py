"""CPU oracle for the power-spectrum hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this module, and only as the checker.  The product path (``jax_powspec_amd``)
never imports it and fails loudly when its HIP library is missing.

This is a NumPy FP64 restatement of the reference's algorithm, function by function:

* ``cic_mas_vec`` / ``cic_mas``  <- src/mas.py:88-153 / src/mas.py:5-87
  (reference stencil, including the (ix, iyp, izp) corner weight of src/mas.py:149 / :76,
  truncating ``int32`` cell index, unwrapped base index, JAX scatter drop rules).
  The position -> (cell, fraction) step is evaluated in float32 exactly as the reference
  evaluates it on float32 inputs; deposit weights and accumulation are FP64.
* ``cic_mesh`` / ``tsc_mesh``    <- textbook CIC / TSC on the same grid convention (grid
  point i at i*H), periodic.  No reference counterpart (extension).
* ``contrast``                   <- caller convention tests/lognormal.py:79-80.
* ``powspec_vec``                <- src/correlations.py:7-56 with jnp.histogram semantics
  (searchsorted side='right', last edge closed, out-of-range dropped).
* ``powspec_vec_fundamental``    <- src/correlations.py:60-117 (k output: mode-averaged k,
  which deviates from the reference's ``.set`` at :100).
* ``subtract_shot_noise``        <- tests/correlations.py:36,83 (monopole only).
* ``xi_grid``                    <- src/correlations.py:134-155: irfftn of
  |rfftn(delta) * correction|^2 (the separation grid).
* ``xi_vec``                     <- src/correlations.py:119-187: histogram of |s| (grid
  units) over the full n^3 grid with edges s_edges n / L, mu = s_z / |s| (NaN at s = 0),
  Nmodes -> inf in empty bins (xi 0 there), xi_l = (2l+1) sum / N / n^3, r = edge midpoints
  times L / n.  kedges = s_edges n / L in FP64 (the reference forms kF s n / L / kF in
  float32).
* ``bispec``                     <- src/correlations.py:334-458: shells
  k_min <= |k| < k_max, k_min/max = k_i/kF -/+ 1 (grid units), of k1, k2 and
  k3(theta) = |k2 (sin t, cos t) + (0, k1)|; delta_i = irfftn(shell delta_k W),
  I_i = irfftn(shell); P_i = sum delta_i^2 / sum I_i^2 (L/n^2)^3,
  B = sum delta1 delta2 delta3 / sum I1 I2 I3 (L^2/n^3)^3, Q = B / (P1P2 + P1P3 + P2P3).
  Shell membership is decided on integer m by FP64 thresholds (the reference compares
  float32 |k| with float32 bounds).
* ``xi_vec_fundamental``         <- src/correlations.py:191-260: slots int(|s|) = isqrt(m),
  slot 0 dropped, r = mean |s| times L / n, 0 / 0 = NaN in empty slots.

How each function is pinned (independently of this file) is listed in tests/test_oracle.py.
The jax/jaxlib arithmetic the reference delegates to (XLA scatter, DUCC FFT, jnp.histogram)
is unvendored and unpinned in the reference; its published semantics are restated here.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32
F64 = np.float64

# wrap modes of the reference's two assigners
WRAP_VEC_TRUE = 0    # src/mas.py:119-122  ixp = (ix + 1 + n) % n
WRAP_VEC_FALSE = 1   # src/mas.py:123-134  ixp = 0 if ix + 1 >= n (ddx is never zeroed)
WRAP_SCAN_TRUE = 2   # src/mas.py:20-28    ixp = ixp - n if ixp >= n
WRAP_SCAN_FALSE = 3  # src/mas.py:30-38    ixp, ddx = 0, 0 if ixp >= n (mdx keeps 1 - ddx)


# ----------------------------------------------------------------------------------------
# mass assignment
# ----------------------------------------------------------------------------------------
def scaled_positions(x, xmin, box_size, n_bins):
    """float32 (x - xmin) * (1 / (box_size / n_bins)), src/mas.py:100-105."""
    bin_size = F32(box_size) / F32(n_bins)
    inv = F32(1.0) / bin_size
    return (np.asarray(x, dtype=F32) - F32(xmin)) * inv


def _normalise(idx, n):
    """JAX indexing: negative index i -> i + n; anything still outside [0, n) is dropped."""
    idx = np.where(idx < 0, idx + n, idx)
    return idx, (idx >= 0) & (idx < n)


def _plus_index(ix, dd, n, mode):
    """+1 neighbour and (possibly zeroed) fraction for the four reference wrap modes."""
    ixp = ix + 1
    if mode == WRAP_VEC_TRUE:
        ixp = np.mod(ixp + n, n)                     # floor mod, jnp '%' semantics
    elif mode == WRAP_VEC_FALSE:
        ixp = np.where(ixp >= n, 0, ixp)             # src/mas.py:125; :126 tests the new ixp
    elif mode == WRAP_SCAN_TRUE:
        ixp = np.where(ixp >= n, ixp - n, ixp)       # src/mas.py:24
    elif mode == WRAP_SCAN_FALSE:
        over = ixp >= n                              # src/mas.py:34
        ixp = np.where(over, 0, ixp)
        dd = np.where(over, F32(0.0), dd)
    else:
        raise ValueError(mode)
    return ixp, dd


def cic_ref_deposits(x, y, z, w, xmin, ymin, zmin, box_size, n_bins, mode):
    """The 8 (index, weight) deposits per particle of the reference stencil.

    Returns a list of 8 tuples (ix, iy, iz, valid, weight_fp64) in the order of
    src/mas.py:142-151.  Index/fraction arithmetic is float32 (as in the reference);
    weights are FP64 products of the float32 factors.
    """
    n = int(n_bins)
    pos = [scaled_positions(c, cmin, box_size, n) for c, cmin in ((x, xmin), (y, ymin), (z, zmin))]
    # non-finite or |pos| >= 2^30 (where int32 truncation is undefined) is dropped
    ok = _finite_cells(pos[0]) & _finite_cells(pos[1]) & _finite_cells(pos[2])
    pos = [np.where(ok, p, F32(0.0)) for p in pos]
    base, dd, md = [], [], []
    for p in pos:
        i = p.astype(np.int64)                        # jnp.int32(): truncation toward 0
        d = (p - i.astype(F32)).astype(F32)           # src/mas.py:111-113
        base.append(i)
        dd.append(d)
        md.append((F32(1.0) - d).astype(F32))         # src/mas.py:138-140 (:56-58)
    plus = []
    dd_w = []
    for a in range(3):
        ip, d = _plus_index(base[a], dd[a], n, mode)
        plus.append(ip)
        dd_w.append(d)
    w64 = np.asarray(w, dtype=F32).astype(F64)
    D = [d.astype(F64) for d in dd_w]
    M = [m.astype(F64) for m in md]
    X = {0: base[0], 1: plus[0]}
    Y = {0: base[1], 1: plus[1]}
    Z = {0: base[2], 1: plus[2]}
    # (sel_x, sel_y, sel_z, wx, wy, wz) in the order of src/mas.py:142-151
    spec = [
        (0, 0, 0, M[0], M[1], M[2]),
        (1, 0, 0, D[0], M[1], M[2]),
        (0, 1, 0, M[0], D[1], M[2]),
        (0, 0, 1, M[0], M[1], D[2]),
        (1, 1, 0, D[0], D[1], M[2]),
        (1, 0, 1, D[0], M[1], D[2]),
        (0, 1, 1, M[0], M[1], D[2]),   # src/mas.py:149 (and :76): mdx*mdy*ddz, not mdx*ddy*ddz
        (1, 1, 1, D[0], D[1], D[2]),
    ]
    out = []
    for sx, sy, sz, wx, wy, wz in spec:
        ix, vx = _normalise(X[sx], n)
        iy, vy = _normalise(Y[sy], n)
        iz, vz = _normalise(Z[sz], n)
        out.append((ix, iy, iz, vx & vy & vz & ok, wx * wy * wz * w64))
    return out


def _accumulate(mesh, n, deposits):
    """Scatter-add every valid deposit (duplicates accumulate), FP64."""
    flat = mesh.reshape(-1)
    for ix, iy, iz, valid, wt in deposits:
        sel = valid
        lin = (ix[sel] * n + iy[sel]) * n + iz[sel]
        flat += np.bincount(lin, weights=wt[sel], minlength=n ** 3)
    return mesh


def cic_mas_vec(delta, x, y, z, w, n_part, xmin, ymin, zmin, box_size, n_bins, wrap):
    """FP64 restatement of src/mas.py:88-153 (n_part is unused there, :91)."""
    n = int(n_bins)
    mesh = np.array(delta, dtype=F64, copy=True).reshape(n, n, n)
    mode = WRAP_VEC_TRUE if wrap else WRAP_VEC_FALSE
    return _accumulate(mesh, n, cic_ref_deposits(x, y, z, w, xmin, ymin, zmin, box_size, n, mode))


def cic_mas(delta, x, y, z, w, n_part, xmin, ymin, zmin, box_size, n_bins, wrap):
    """FP64 restatement of src/mas.py:5-87 (the lax.scan assigner)."""
    n = int(n_bins)
    mesh = np.array(delta, dtype=F64, copy=True).reshape(n, n, n)
    mode = WRAP_SCAN_TRUE if wrap else WRAP_SCAN_FALSE
    return _accumulate(mesh, n, cic_ref_deposits(x, y, z, w, xmin, ymin, zmin, box_size, n, mode))


def _finite_cells(p):
    ok = np.isfinite(p) & (np.abs(p) < F32(2.0 ** 30))
    return ok


def cic_mesh(x, y, z, w, xmin, ymin, zmin, box_size, n_bins, delta=None):
    """Textbook (mass-conserving) periodic CIC; grid point i at i*H.  Extension."""
    n = int(n_bins)
    mesh = np.zeros((n, n, n), F64) if delta is None else np.array(delta, F64, copy=True)
    pos = [scaled_positions(c, cmin, box_size, n) for c, cmin in ((x, xmin), (y, ymin), (z, zmin))]
    ok = _finite_cells(pos[0]) & _finite_cells(pos[1]) & _finite_cells(pos[2])
    pos = [p[ok] for p in pos]
    w64 = np.asarray(w, dtype=F32)[ok].astype(F64)
    i0, d = [], []
    for p in pos:
        f = np.floor(p).astype(np.int64)
        i0.append(f)
        d.append(p.astype(F64) - f)                  # exact: p - floor(p) is representable
    deps = []
    for sx in (0, 1):
        for sy in (0, 1):
            for sz in (0, 1):
                wt = w64.copy()
                for a, s in enumerate((sx, sy, sz)):
                    wt *= d[a] if s else (1.0 - d[a])
                deps.append((np.mod(i0[0] + sx, n), np.mod(i0[1] + sy, n), np.mod(i0[2] + sz, n),
                             np.ones(wt.shape, bool), wt))
    return _accumulate(mesh, n, deps)


def tsc_mesh(x, y, z, w, xmin, ymin, zmin, box_size, n_bins, delta=None):
    """Textbook periodic TSC; nearest grid point i0 = floor(pos + 0.5) (float32), d = pos - i0.

    W(i0) = 3/4 - d^2, W(i0 +- 1) = (1/2)(1/2 -+ d)^2 ... written here as
    W(i0-1) = 0.5 (0.5 - d)^2, W(i0+1) = 0.5 (0.5 + d)^2.  Extension (no reference).
    """
    n = int(n_bins)
    mesh = np.zeros((n, n, n), F64) if delta is None else np.array(delta, F64, copy=True)
    pos = [scaled_positions(c, cmin, box_size, n) for c, cmin in ((x, xmin), (y, ymin), (z, zmin))]
    ok = _finite_cells(pos[0]) & _finite_cells(pos[1]) & _finite_cells(pos[2])
    pos = [p[ok] for p in pos]
    w64 = np.asarray(w, dtype=F32)[ok].astype(F64)
    i0, wts = [], []
    for p in pos:
        c = np.floor((p + F32(0.5)).astype(F32)).astype(np.int64)
        d = p.astype(F64) - c
        i0.append(c)
        wts.append((0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2))
    deps = []
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                wt = w64 * wts[0][sx + 1] * wts[1][sy + 1] * wts[2][sz + 1]
                deps.append((np.mod(i0[0] + sx, n), np.mod(i0[1] + sy, n), np.mod(i0[2] + sz, n),
                             np.ones(wt.shape, bool), wt))
    return _accumulate(mesh, n, deps)


def assign(stencil, x, y, z, w, xmin, ymin, zmin, box_size, n_bins, delta=None, wrap=True):
    """Dispatch by stencil name: 'CIC_REF' (reference cic_mas_vec), 'CIC', 'TSC'."""
    n = int(n_bins)
    if w is None:
        w = np.ones(np.asarray(x).shape, F32)
    if stencil == "CIC_REF":
        base = np.zeros((n, n, n), F64) if delta is None else delta
        return cic_mas_vec(base, x, y, z, w, None, xmin, ymin, zmin, box_size, n, wrap)
    if stencil == "CIC":
        return cic_mesh(x, y, z, w, xmin, ymin, zmin, box_size, n, delta)
    if stencil == "TSC":
        return tsc_mesh(x, y, z, w, xmin, ymin, zmin, box_size, n, delta)
    raise ValueError(stencil)


def contrast(mesh):
    """delta / mean(delta) - 1, tests/lognormal.py:79-80 (FP64)."""
    mesh = np.asarray(mesh, F64)
    return mesh / mesh.mean() - 1.0


# ----------------------------------------------------------------------------------------
# power spectrum
# ----------------------------------------------------------------------------------------
def wavenumbers(n):
    """Integer wavenumbers ki of src/correlations.py:23-25 (Nyquist stays +n/2)."""
    r = np.arange(n)
    return np.where(r > n // 2, r - n, r)


def half_spectrum_grids(n):
    ki = wavenumbers(n)
    mid = n // 2
    kx = ki[:, None, None]
    ky = ki[None, :, None]
    kz = ki[None, None, : mid + 1]
    return kx, ky, kz


def window_correction(n, p):
    """prod_a (1 / sinc(k_a / n))^p over the half-spectrum, src/correlations.py:20-21,32."""
    kx, ky, kz = half_spectrum_grids(n)
    c = lambda k: (1.0 / np.sinc(k / n)) ** p   # np.sinc(t) = sin(pi t)/(pi t); x/pi = k/n
    return c(kx) * c(ky) * c(kz)


def jnp_histogram_slots(k, edges):
    """Slot index (or -1 = dropped) under jnp.histogram's rules.

    bin_idx = searchsorted(edges, k, side='right'); k == edges[-1] -> len(edges)-1;
    the result is scattered into zeros(len(edges)) and [1:] is kept, so idx 0 and
    idx len(edges) (out of bounds) are dropped.
    """
    edges = np.asarray(edges, F64)
    nb = edges.size - 1
    idx = np.searchsorted(edges, k, side="right")
    idx = np.where(k == edges[-1], nb, idx)
    slot = idx - 1
    return np.where((idx >= 1) & (idx <= nb), slot, -1)


def legendre_weights(kz2, m):
    """L2(mu), L4(mu) for mu^2 = kz^2 / m from exact integer numerators (FP64).

    (3 mu^2 - 1)/2 = (3 kz^2 - m) / (2 m) and (35 mu^4 - 30 mu^2 + 3)/8 =
    (35 kz^4 - 30 kz^2 m + 3 m^2) / (8 m^2): one rounding each, and exactly 0 where the
    numerator is 0 (e.g. 3 kz^2 = m).  mu = 0 at m = 0 (src/correlations.py:37).
    """
    kz2 = np.asarray(kz2, np.int64)
    m = np.asarray(m, np.int64)
    mm = np.where(m == 0, 1, m).astype(np.float64)
    L2 = np.where(m == 0, -0.5, (3 * kz2 - m).astype(np.float64) / (2.0 * mm))
    L4 = np.where(m == 0, 0.375, (35 * kz2 * kz2 - 30 * kz2 * m + 3 * m * m).astype(np.float64) / (8.0 * mm * mm))
    return L2, L4


def _bin_sums(slot, nb, d2, kz2, m, kk):
    sel = slot >= 0
    s = slot[sel]
    P = d2[sel]
    L2, L4 = legendre_weights(kz2[sel], m[sel])
    S0 = np.bincount(s, weights=P, minlength=nb)[:nb]
    S2 = np.bincount(s, weights=P * L2, minlength=nb)[:nb]
    S4 = np.bincount(s, weights=P * L4, minlength=nb)[:nb]
    A2 = np.bincount(s, weights=P * np.abs(L2), minlength=nb)[:nb]
    A4 = np.bincount(s, weights=P * np.abs(L4), minlength=nb)[:nb]
    cnt = np.bincount(s, minlength=nb)[:nb].astype(np.int64)
    ksum = np.bincount(s, weights=kk[sel], minlength=nb)[:nb]
    return S0, S2, S4, cnt, ksum, A2, A4


def spectrum_power(delta, p=2, zero_dc=False):
    """(|delta_k|^2 * correction^2, k, kz^2, m) over the stored half spectrum (FP64)."""
    delta = np.asarray(delta, F64)
    n = delta.shape[0]
    dk = np.fft.rfftn(delta, axes=(0, 1, 2))            # src/correlations.py:22
    if zero_dc:
        dk[0, 0, 0] = 0.0
    dk = dk * window_correction(n, p)                   # :39
    d2 = (dk * dk.conj()).real                          # :40
    kx, ky, kz = half_spectrum_grids(n)
    m = (kx * kx + ky * ky + kz * kz).astype(np.int64)
    k = np.sqrt(m.astype(F64))                          # :34
    kz2 = np.broadcast_to(kz * kz, m.shape).astype(np.int64)      # mu^2 = kz^2 / k^2, :36-38
    return d2.reshape(-1), k.reshape(-1), kz2.reshape(-1), m.reshape(-1)


def _normalise_pk(S0, S2, S4, cnt, box_size, n):
    norm = (box_size / n ** 2) ** 3                     # :50-52
    with np.errstate(invalid="ignore", divide="ignore"):
        Pk = np.stack([S0 / cnt * norm, S2 / cnt * 5.0 * norm, S4 / cnt * 9.0 * norm], axis=1)
    return Pk


def powspec_vec(delta, box_size, k_edges, p=2, zero_dc=False, full=False):
    """FP64 restatement of src/correlations.py:7-56.

    Returns (k3D, Pk3D[nb,3], Nmodes3D) as the reference; Pk3D is FP64 here (the reference
    casts to float32 at :49).  Empty bins give NaN (0/0), as in the reference.
    With full=True also returns a dict with FP64 sums, int64 counts, the per-bin
    same-bin absolute scales A_l = sum |L_l| P norm (2l+1) / N and the mode-averaged k.
    """
    delta = np.asarray(delta, F64)
    n = delta.shape[0]
    kF = 2.0 * np.pi / box_size                          # :12
    d2, k, kz2, m = spectrum_power(delta, p, zero_dc)
    kedges = np.asarray(k_edges, F64) / kF               # :42
    nb = kedges.size - 1
    slot = jnp_histogram_slots(k, kedges)
    S0, S2, S4, cnt, ksum, A2, A4 = _bin_sums(slot, nb, d2, kz2, m, k)
    Pk = _normalise_pk(S0, S2, S4, cnt.astype(F64), box_size, n)
    k3D = 0.5 * (kedges[1:] + kedges[:-1]) * kF          # :54
    Nmodes = cnt.astype(F64)
    if not full:
        return k3D, Pk, Nmodes
    norm = (box_size / n ** 2) ** 3
    with np.errstate(invalid="ignore", divide="ignore"):
        scale = np.stack([S0 / cnt * norm, A2 / cnt * 5 * norm, A4 / cnt * 9 * norm], axis=1)
        kmean = ksum / cnt * kF
    extra = dict(S=np.stack([S0, S2, S4], 1), counts=cnt, scale=scale, kmean=kmean, slot=slot)
    return k3D, Pk, Nmodes, extra


def powspec_vec_fundamental(delta, box_size, p=2, zero_dc=False, full=False):
    """FP64 restatement of src/correlations.py:60-117.

    Bins are int32(k) = isqrt(m) (exact), kmax = int32(sqrt(3 middle^2)), slot 0 (DC)
    dropped (:106-109).  k3D is the mode-averaged k (the reference's :100 uses .set, whose
    value is the last writer; that value cannot be pinned and is not reproduced).
    """
    delta = np.asarray(delta, F64)
    n = delta.shape[0]
    mid = n // 2
    kF = 2.0 * np.pi / box_size
    kmax = math.isqrt(3 * mid * mid)                     # :69 (3*mid^2 is never a square)
    d2, k, kz2, m = spectrum_power(delta, p, zero_dc)
    # :87 int32(k); floor of the correctly rounded FP64 sqrt equals isqrt(m) for m < 2**50
    kindex = np.floor(k).astype(np.int64)
    slot = kindex - 1                                    # drop slot 0 (:106-109)
    nb = kmax
    S0, S2, S4, cnt, ksum, A2, A4 = _bin_sums(slot, nb, d2, kz2, m, k)
    Pk = _normalise_pk(S0, S2, S4, cnt.astype(F64), box_size, n)
    with np.errstate(invalid="ignore", divide="ignore"):
        k3D = ksum / cnt * kF
    if not full:
        return k3D, Pk, cnt.astype(F64)
    norm = (box_size / n ** 2) ** 3
    with np.errstate(invalid="ignore", divide="ignore"):
        scale = np.stack([S0 / cnt * norm, A2 / cnt * 5 * norm, A4 / cnt * 9 * norm], axis=1)
    extra = dict(S=np.stack([S0, S2, S4], 1), counts=cnt, scale=scale, kmean=k3D, slot=slot)
    return k3D, Pk, cnt.astype(F64), extra


def subtract_shot_noise(Pk, shot_noise):
    """P_0 - shot noise (monopole only), tests/correlations.py:83 / tests/create_lognormal.py:87."""
    out = np.array(Pk, F64, copy=True)
    out[:, 0] -= shot_noise
    return out


def estimate(x, y, z, w, box_size, n, k_edges=None, stencil="CIC_REF", p=2, shot_noise=0.0,
             xmin=0.0, ymin=0.0, zmin=0.0, full=True):
    """Whole path: assign -> contrast -> P_l (user edges or k_f bins) -> shot noise."""
    mesh = assign(stencil, x, y, z, w, xmin, ymin, zmin, box_size, n)
    delta = contrast(mesh)
    if k_edges is None:
        res = powspec_vec_fundamental(delta, box_size, p=p, zero_dc=True, full=full)
    else:
        res = powspec_vec(delta, box_size, k_edges, p=p, zero_dc=True, full=full)
    k3D, Pk = res[0], res[1]
    Pk = subtract_shot_noise(Pk, shot_noise)
    return (k3D, Pk) + tuple(res[2:]) + (mesh,)


# ----------------------------------------------------------------------------------------
# reverse-mode derivative of estimate() (checker of the hand-written adjoint)
# ----------------------------------------------------------------------------------------
def corner_stencil(stencil, pos, n, mode=WRAP_VEC_TRUE):
    """Corners of every particle with unit-weight W and dW/dpos per axis (FP64).

    pos: three arrays of scaled positions (x - xmin) * n / box.  Same corners, wrap and
    drop rules as cic_ref_deposits / cic_mesh / tsc_mesh.  Returns a list of
    (flat_index, valid, W, (dWx, dWy, dWz)).
    """
    pos = [np.asarray(p, F64) for p in pos]
    ok = np.ones(pos[0].shape, bool)
    for p in pos:
        ok &= np.isfinite(p) & (np.abs(p) < 2.0 ** 30)
    pos = [np.where(ok, p, 0.0) for p in pos]
    axes = []   # per axis: list of (index, valid, weight, slope) options
    if stencil == "CIC_REF":
        for p in pos:
            i = np.trunc(p).astype(np.int64)
            d = p - i
            ip, dz = _plus_index(i, d, n, mode)
            slope_p = np.where(dz == d, 1.0, 0.0) if mode == WRAP_SCAN_FALSE else np.ones_like(d)
            ib, vb = _normalise(i, n)
            ipn, vp = _normalise(ip, n)
            axes.append(((ib, vb, 1.0 - d, -np.ones_like(d)), (ipn, vp, dz, slope_p)))
        # src/mas.py:142-151 selections; the :149 corner takes md_y with the index iyp
        spec = [(0, 0, 0, 0, 0, 0), (1, 0, 0, 1, 0, 0), (0, 1, 0, 0, 1, 0), (0, 0, 1, 0, 0, 1),
                (1, 1, 0, 1, 1, 0), (1, 0, 1, 1, 0, 1), (0, 1, 1, 0, 0, 1), (1, 1, 1, 1, 1, 1)]
        out = []
        for ix, iy, iz, wx, wy, wz in spec:
            X, Y, Z = axes[0][ix], axes[1][iy], axes[2][iz]
            WX, WY, WZ = axes[0][wx], axes[1][wy], axes[2][wz]
            flat = (X[0] * n + Y[0]) * n + Z[0]
            valid = X[1] & Y[1] & Z[1] & ok
            W = WX[2] * WY[2] * WZ[2]
            dW = (WX[3] * WY[2] * WZ[2], WX[2] * WY[3] * WZ[2], WX[2] * WY[2] * WZ[3])
            out.append((flat, valid, W, dW))
        return out
    for p in pos:
        if stencil == "CIC":
            f = np.floor(p)
            d = p - f
            i = f.astype(np.int64)
            axes.append([(np.mod(i, n), 1.0 - d, -np.ones_like(d)), (np.mod(i + 1, n), d, np.ones_like(d))])
        elif stencil == "TSC":
            c = np.floor(p + 0.5)
            d = p - c
            i = c.astype(np.int64)
            axes.append([(np.mod(i - 1, n), 0.5 * (0.5 - d) ** 2, -(0.5 - d)),
                         (np.mod(i, n), 0.75 - d * d, -2.0 * d),
                         (np.mod(i + 1, n), 0.5 * (0.5 + d) ** 2, 0.5 + d)])
        else:
            raise ValueError(stencil)
    out = []
    for X in axes[0]:
        for Y in axes[1]:
            for Z in axes[2]:
                flat = (X[0] * n + Y[0]) * n + Z[0]
                out.append((flat, ok, X[1] * Y[1] * Z[1], (X[2] * Y[1] * Z[1], X[1] * Y[2] * Z[1], X[1] * Y[1] * Z[2])))
    return out


def estimate_fp64(x, y, z, w, box_size, n, k_edges, stencil="CIC_REF", p=2, mode=WRAP_VEC_TRUE):
    """estimate() with FP64 positions (no float32 step): the function whose finite
    differences pin estimate_backward.  Returns (Pk[nb, 3] before shot noise, counts, mesh)."""
    n = int(n)
    inv = n / float(box_size)
    pos = [np.asarray(c, F64) * inv for c in (x, y, z)]
    w = np.ones(pos[0].shape) if w is None else np.asarray(w, F64)
    rho = np.zeros(n ** 3)
    for flat, valid, W, _ in corner_stencil(stencil, pos, n, mode):
        rho += np.bincount(flat[valid], weights=(w * W)[valid], minlength=n ** 3)
    mesh = rho.reshape(n, n, n)
    delta = contrast(mesh)
    if k_edges is None:
        res = powspec_vec_fundamental(delta, box_size, p=p, zero_dc=True, full=True)
    else:
        res = powspec_vec(delta, box_size, k_edges, p=p, zero_dc=True, full=True)
    return res[1], res[3]["counts"], mesh


def estimate_backward(x, y, z, w, box_size, n, k_edges, grad_pk, stencil="CIC_REF", p=2, mode=WRAP_VEC_TRUE):
    """d(sum grad_pk * Pk)/d(positions, weights) of estimate_fp64, by the chain rule.

    P_lb = c_lb sum_{k in b} W_k L_l |delta_k|^2 (half spectrum), delta = rho n^3 / M - 1,
    rho = sum_i w_i W(x - p_i).  Returns (grad_pos[N, 3], grad_w[N]).
    """
    n = int(n)
    inv = n / float(box_size)
    pos = [np.asarray(c, F64) * inv for c in (x, y, z)]
    w = np.ones(pos[0].shape) if w is None else np.asarray(w, F64)
    corners = corner_stencil(stencil, pos, n, mode)
    rho = np.zeros(n ** 3)
    for flat, valid, W, _ in corners:
        rho += np.bincount(flat[valid], weights=(w * W)[valid], minlength=n ** 3)
    M = rho.sum()
    S = n ** 3 / M
    delta = (rho * S - 1.0).reshape(n, n, n)
    dk = np.fft.rfftn(delta)
    dk[0, 0, 0] = 0.0
    kx, ky, kz = half_spectrum_grids(n)
    m = (kx * kx + ky * ky + kz * kz).astype(np.int64)
    kz2 = np.broadcast_to(kz * kz, m.shape).astype(np.int64)
    if k_edges is None:
        slot = np.floor(np.sqrt(m.astype(F64))).astype(np.int64) - 1
        nb = math.isqrt(3 * (n // 2) ** 2)
        slot = np.where(slot < nb, slot, -1)
    else:
        kedges = np.asarray(k_edges, F64) / (2.0 * np.pi / box_size)
        nb = kedges.size - 1
        slot = jnp_histogram_slots(np.sqrt(m.astype(F64)), kedges)
    counts = np.bincount(slot[slot >= 0], minlength=nb)[:nb]
    norm = (box_size / n ** 2) ** 3
    g = np.asarray(grad_pk, F64).reshape(nb, 3)
    with np.errstate(invalid="ignore", divide="ignore"):
        coef = np.where(counts[:, None] > 0, g * norm * np.array([1.0, 5.0, 9.0]) / counts[:, None], 0.0)
    L2, L4 = legendre_weights(kz2, m)
    s = np.where(slot >= 0, slot, 0)
    G = window_correction(n, 2 * p) * (coef[s, 0] + coef[s, 1] * L2 + coef[s, 2] * L4)
    G = np.where((slot >= 0) & (m > 0), G, 0.0)
    A = 2.0 * G * dk
    # dLoss/d delta(x) = Re sum_{half} A_k e^{+i k.x}
    full = np.zeros((n, n, n), complex)
    full[:, :, : n // 2 + 1] = A
    gx = (np.fft.ifftn(full) * n ** 3).real.reshape(-1)
    gbar = float(np.dot(gx, rho)) / M
    ga = np.zeros((pos[0].size, 3))
    gw = np.zeros(pos[0].size)
    for flat, valid, W, dW in corners:
        gv = np.where(valid, gx[np.where(valid, flat, 0)], 0.0)
        vm = valid.astype(F64)
        for ax in range(3):
            ga[:, ax] += (gv - gbar) * dW[ax] * vm
        gw += (gv - gbar) * W * vm
    grad_pos = S * inv * w[:, None] * ga
    grad_w = S * gw
    return grad_pos, grad_w


# ------------------------------------------------------------------------------------
# configuration-space multipoles (src/correlations.py:119-260)
# ------------------------------------------------------------------------------------
def real_grids(n):
    """Signed separations (ki of src/correlations.py:139-141) on the full n^3 grid."""
    ki = wavenumbers(n)
    return ki[:, None, None], ki[None, :, None], ki[None, None, :]


def xi_grid(delta, p=2):
    """irfftn(|rfftn(delta) correction|^2) in FP64 (src/correlations.py:134-155)."""
    delta = np.asarray(delta, F64)
    n = delta.shape[0]
    dk = np.fft.rfftn(delta, axes=(0, 1, 2)) * window_correction(n, p)
    return np.fft.irfftn((dk * dk.conj()).real, s=(n, n, n), axes=(0, 1, 2))


def _xi_sums(g, slot, nb, m, sz2):
    sel = slot >= 0
    s = slot[sel]
    v = g.reshape(-1)[sel]
    mm = m[sel]
    z2 = sz2[sel]
    with np.errstate(invalid="ignore", divide="ignore"):
        L2 = (3 * z2 - mm).astype(F64) / (2.0 * mm)             # NaN at s = 0 (mu = 0/0)
        L4 = (35 * z2 * z2 - 30 * z2 * mm + 3 * mm * mm).astype(F64) / (8.0 * mm.astype(F64) ** 2)
    S = np.stack([np.bincount(s, weights=v * w, minlength=nb)[:nb] for w in (1.0, L2, L4)], axis=1)
    with np.errstate(invalid="ignore"):
        A = np.stack([np.bincount(s, weights=np.abs(v * w), minlength=nb)[:nb] for w in (1.0, L2, L4)], axis=1)
    cnt = np.bincount(s, minlength=nb)[:nb].astype(np.int64)
    ksum = np.bincount(s, weights=np.sqrt(mm.astype(F64)), minlength=nb)[:nb]
    return S, A, cnt, ksum


def _xi_full_grid(n):
    sx, sy, sz = real_grids(n)
    m = (sx * sx + sy * sy + sz * sz).astype(np.int64)
    sz2 = np.broadcast_to(sz * sz, m.shape).astype(np.int64)
    return m.reshape(-1), sz2.reshape(-1)


def xi_vec(delta, box_size, s_edges, p=2):
    """FP64 restatement of src/correlations.py:119-187.

    Returns (r3D, xi3D[nb, 3], Nmodes3D, extra): Nmodes3D is inf in empty bins, where xi is
    0; extra holds int64 counts and the same-bin absolute scale A[nb, 3] =
    (2l+1) sum |xi_cell L_l| / (N n^3) used as the parity tolerance unit."""
    n = np.asarray(delta).shape[0]
    g = xi_grid(delta, p)
    m, sz2 = _xi_full_grid(n)
    kedges = np.asarray(s_edges, F64) * n / float(box_size)
    slot = jnp_histogram_slots(np.sqrt(m.astype(F64)), kedges)
    nb = kedges.size - 1
    S, A, cnt, _ = _xi_sums(g, slot, nb, m, sz2)
    N = np.where(cnt == 0, np.inf, cnt.astype(F64))
    f = np.array([1.0, 5.0, 9.0]) / float(n) ** 3
    xi = S / N[:, None] * f
    r = 0.5 * (kedges[1:] + kedges[:-1]) * (float(box_size) / n)
    with np.errstate(invalid="ignore"):
        scale = A / N[:, None] * f
    return r, xi, N, {"counts": cnt, "scale": scale}


def xi_vec_fundamental(delta, box_size, p=2):
    """FP64 restatement of src/correlations.py:191-260 (slots j = isqrt(m), j >= 1)."""
    n = np.asarray(delta).shape[0]
    g = xi_grid(delta, p)
    m, sz2 = _xi_full_grid(n)
    kmax = int(np.sqrt(3 * (n // 2) ** 2))
    j = np.array([int(v) for v in np.sqrt(np.arange(m.max() + 1, dtype=F64))])[m]   # int32(|s|)
    slot = np.where(j >= 1, j - 1, -1)
    S, A, cnt, ksum = _xi_sums(g, slot, kmax, m, sz2)
    f = np.array([1.0, 5.0, 9.0]) / float(n) ** 3
    with np.errstate(invalid="ignore", divide="ignore"):
        xi = S / cnt[:, None] * f
        r = ksum / cnt * (float(box_size) / n)
        scale = A / cnt[:, None] * f
    return r, xi, cnt.astype(F64), {"counts": cnt, "scale": scale}


# ------------------------------------------------------------------------------------
# bispectrum (src/correlations.py:334-458)
# ------------------------------------------------------------------------------------
def bispec_shells(box_size, k1, k2, theta):
    """k_all = (k1, k2, k3(theta)) and shell bounds in grid units (:352-359)."""
    kF = 2.0 * np.pi / float(box_size)
    theta = np.asarray(theta, F64)
    k3 = np.sqrt((k2 * np.sin(theta)) ** 2 + (k2 * np.cos(theta) + k1) ** 2)
    k_all = np.concatenate([[k1, k2], k3]).astype(F64)
    return k_all, (k_all - kF) / kF, (k_all + kF) / kF


def bispec(delta, box_size, k1, k2, theta, p=1):
    """FP64 restatement of src/correlations.py:334-458: (k_all, Pk[bins+2], theta, B, Q) and
    extras (per-shell sums and the scale of the triple-product sums).  p is the window row:
    delta_k is multiplied by prod (x / sin x)^(2p), so p = 1 is the reference's CIC
    correction of index 2 (:362-373)."""
    delta = np.asarray(delta, F64)
    n = delta.shape[0]
    k_all, kmin, kmax = bispec_shells(box_size, k1, k2, theta)
    dk = np.fft.rfftn(delta, axes=(0, 1, 2)) * window_correction(n, 2 * p)  # :362-373
    kx, ky, kz = half_spectrum_grids(n)
    m = (kx * kx + ky * ky + kz * kz).astype(np.int64)
    k = np.sqrt(m.astype(F64))
    fields = []
    for lo, hi in zip(kmin, kmax):
        mask = (k >= lo) & (k < hi)                                          # :377
        d = np.fft.irfftn(np.where(mask, dk, 0.0), s=(n, n, n), axes=(0, 1, 2))
        I = np.fft.irfftn(mask.astype(F64), s=(n, n, n), axes=(0, 1, 2))
        fields.append((d, I))
    norm_p = (box_size / n ** 2) ** 3
    norm_b = (box_size ** 2 / n ** 3) ** 3
    Pk = np.array([(d * d).sum() / (I * I).sum() * norm_p for d, I in fields])
    d1, I1 = fields[0]
    d2, I2 = fields[1]
    B, Q, trip, tri, scale = [], [], [], [], []
    for i in range(2, len(fields)):
        d3, I3 = fields[i]
        t = (d1 * d2 * d3).sum()
        c = (I1 * I2 * I3).sum()
        b = t / c * norm_b
        B.append(b)
        Q.append(b / (Pk[0] * Pk[1] + Pk[0] * Pk[i] + Pk[1] * Pk[i]))
        trip.append(t)
        tri.append(c)
        scale.append(np.abs(d1 * d2 * d3).sum() / abs(c) * norm_b)
    extra = {"sum_d2": np.array([(d * d).sum() for d, _ in fields]),
             "sum_I2": np.array([(I * I).sum() for _, I in fields]),
             "triples": np.array(trip), "triangles": np.array(tri), "scale_B": np.array(scale)}
    return k_all, Pk, np.asarray(theta, F64), np.array(B), np.array(Q), extra
