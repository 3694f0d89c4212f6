"""CPU oracle for the seeded lognormal mock generator -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this module, and only as the checker.  The product path
(``jax_powspec_amd.gauss_field``) never imports it.

A NumPy restatement of the reference's mock pipeline with the random streams that the
HIP generator defines (counter-based Philox4x32-10; the reference draws from NumPy's
MT19937 and JAX's threefry instead, so individual values cannot coincide with it; the
algorithm around the draws is the reference's):

* ``philox4x32``           <- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the
                              published counter-based generator.
* ``pk_interp``            <- the bisection + linear interpolation of
                              src/gauss_field.py:39-47 (P(|k|) times (n^2/L)^3).
* ``gaussian_field``       <- src/gauss_field.py:5-80: half spectrum (n, n, n/2+1) in the
                              reference's [kxx][kyy][kzz] layout; phase 2 pi u, amplitude
                              sqrt(-ln u) (Rayleigh) or 1, times sqrt(P); on the kz = 0 and
                              kz = n/2 planes the first of (k, -k) in loop order draws and
                              the other is its conjugate (:61-75); self-conjugate modes are
                              real and equal to the amplitude (:74-75); delta(0) = 0 (:78).
                              Draws of mode m = (kxx n + kyy)(n/2+1) + kzz come from the
                              Philox block (m, 0, TAG_FIELD).
* ``irfftn``               <- scipy.fft.irfftn as called at tests/create_lognormal.py:47.
* ``psi_z_spectrum``       <- Zel'dovich displacement along the third axis,
                              psi_z(k) = i k_z delta(k) / k^2 (the nbodykit LogNormalCatalog
                              branch of tests/create_lognormal.py:23-29; the field-level
                              form of src/populate_field.py:57-69 is broken in the
                              reference); the k_z = n/2 plane is set to 0.
* ``poisson``              <- inversion for lambda < 10, Hormann's PTRS transformed
                              rejection (1993) for lambda >= 10; the inversion uniform of
                              cell c from the Philox block (c >> 1, 0, TAG_POISSON) (words
                              0, 1 for even c, 2, 3 for odd c), rejection rounds from
                              blocks (c, j, TAG_PTRS), j = 0, 1, ...
* ``displacements``        <- src/gauss_field.py:83-88: R = 2u - 1,
                              sign(R) (1 - sqrt|R|) times the cell size; particle p draws
                              from the Philox block (p, 0, TAG_DISP).
* ``populate``             <- src/gauss_field.py:90-110 / src/populate_field.py:11-29:
                              lambda = rho (cell volume * density) / mean(rho), Poisson per
                              cell, cell centre + displacement, wrapped into [0, L).  With
                              ``bias`` != 0 rho = exp(bias * field) (the lognormal step of
                              tests/create_lognormal.py:50-51).  Particles come out in
                              C order of the cells (the reference orders cells by
                              decreasing rho, a permutation of the same catalogue).

How each function is pinned (independently of this file) is listed in
tests/test_oracle_mocks.py: Philox's published known-answer vectors; the reference's own
bisection loop; a loop-order emulation of the reference's Hermitian fill; NumPy's FFT
round trip; scipy's Poisson quantiles and chi-square tests against the Poisson pmf; the
triangular displacement law.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
F64 = np.float64
U64 = np.uint64

MASK32 = 0xFFFFFFFF
PHILOX_M0 = 0xD2511F53
PHILOX_M1 = 0xCD9E8D57
PHILOX_W0 = 0x9E3779B9
PHILOX_W1 = 0xBB67AE85

TAG_FIELD = 1
TAG_POISSON = 2
TAG_DISP = 3
TAG_PTRS = 4

INV_KMAX = 100        # inversion search bound (lambda < 10: the tail beyond is < 1e-60)
PTRS_MAX_IT = 64      # rejection rounds before the deterministic fallback floor(lambda)
POISSON_INV_MAX = 10.0
POISSON_LAM_MAX = 2.0 ** 30   # rates are clamped (counts fit int32)


# ------------------------------------------------------------------------------------
# Philox4x32-10
# ------------------------------------------------------------------------------------
def philox4x32(c0, c1, c2, c3, k0, k1, rounds: int = 10):
    """Philox4x32-R on arrays (any broadcastable shapes) of 32-bit words; returns 4 uint64
    arrays holding the 32-bit outputs."""
    c0, c1, c2, c3 = (np.asarray(v, dtype=U64) & U64(MASK32) for v in (c0, c1, c2, c3))
    k0 = np.asarray(k0, dtype=U64) & U64(MASK32)
    k1 = np.asarray(k1, dtype=U64) & U64(MASK32)
    m0, m1, m = U64(PHILOX_M0), U64(PHILOX_M1), U64(MASK32)
    for r in range(rounds):
        if r > 0:
            k0 = (k0 + U64(PHILOX_W0)) & m
            k1 = (k1 + U64(PHILOX_W1)) & m
        p0 = m0 * c0
        p1 = m1 * c2
        hi0, lo0 = p0 >> U64(32), p0 & m
        hi1, lo1 = p1 >> U64(32), p1 & m
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return c0, c1, c2, c3


def seed_key(seed: int):
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return s & MASK32, s >> 32


def uf32(w):
    """((w >> 9) + 1/2) / 2^23: a float32-exact uniform in (0, 1)."""
    return ((np.asarray(w, dtype=U64) >> U64(9)).astype(F64) + 0.5) * 2.0 ** -23


def u52(a, b):
    """((a >> 6) 2^26 + (b >> 6) + 1/2) / 2^52: an FP64-exact uniform in (0, 1)."""
    hi = (np.asarray(a, dtype=U64) >> U64(6)).astype(F64)
    lo = (np.asarray(b, dtype=U64) >> U64(6)).astype(F64)
    return (hi * 67108864.0 + lo + 0.5) * 2.0 ** -52


def block(index, j, tag, seed):
    """The Philox block (index, j, tag) under the seed's key."""
    index = np.asarray(index, dtype=U64)
    k0, k1 = seed_key(seed)
    return philox4x32(index & U64(MASK32), index >> U64(32), j, tag, k0, k1)


# ------------------------------------------------------------------------------------
# Gaussian field
# ------------------------------------------------------------------------------------
def pk_interp(kmod, kf, Pkf):
    """Linear interpolation between the bracketing table entries found by bisection
    (src/gauss_field.py:39-46); kf strictly increasing.  Outside the table the end
    segments extrapolate."""
    kf = np.asarray(kf, dtype=F64)
    Pkf = np.asarray(Pkf, dtype=F64)
    K = len(kf)
    hi = np.clip(np.searchsorted(kf, kmod, side="left"), 1, K - 1)
    lo = hi - 1
    return (Pkf[hi] - Pkf[lo]) / (kf[hi] - kf[lo]) * (kmod - kf[lo]) + Pkf[lo]


def amplitude_table(n, box, kf, Pkf):
    """sqrt(P(k_F sqrt(m)) (n^2/L)^3) for m = 0 .. 3 (n/2)^2 (src/gauss_field.py:36-47, :56)."""
    m = np.arange(3 * (n // 2) ** 2 + 1, dtype=F64)
    kmod = np.sqrt(m) * (2.0 * np.pi / box)
    P = pk_interp(kmod, kf, Pkf) * (float(n) ** 2 / box) ** 3
    with np.errstate(invalid="ignore"):
        return np.sqrt(P)


def mode_grid(n):
    h = n // 2
    kxx = np.arange(n)[:, None, None]
    kyy = np.arange(n)[None, :, None]
    kzz = np.arange(h + 1)[None, None, :]
    return kxx, kyy, kzz


def gaussian_field(grid, kf, Pkf, Rayleigh_sampling, seed, BoxSize):
    """delta_k of src/gauss_field.py:5-80, complex128 (grid, grid, grid//2+1)."""
    n = int(grid)
    h = n // 2
    kxx, kyy, kzz = mode_grid(n)
    kxx, kyy, kzz = np.broadcast_arrays(kxx, kyy, kzz)
    mx, my = (n - kxx) % n, (n - kyy) % n
    a = kxx * n + kyy
    b = mx * n + my
    plane = (kzz == 0) | (kzz == h)
    conj = plane & (b < a)
    selfc = plane & (a == b)
    owner = np.where(conj, b, a).astype(U64)
    mref = owner * U64(h + 1) + kzz.astype(U64)
    w0, w1, _, _ = block(mref, 0, TAG_FIELD, seed)
    u_ph, u_am = uf32(w0), uf32(w1)
    sx = np.where(kxx > h, kxx - n, kxx)
    sy = np.where(kyy > h, kyy - n, kyy)
    m = sx * sx + sy * sy + kzz * kzz
    amp = amplitude_table(n, BoxSize, kf, Pkf)[m]
    if Rayleigh_sampling == 1:
        amp = amp * np.sqrt(-np.log(u_am))
    phase = 2.0 * np.pi * u_ph
    re = amp * np.cos(phase)
    im = amp * np.sin(phase)
    im = np.where(conj, -im, im)
    re = np.where(selfc, amp, re)
    im = np.where(selfc, 0.0, im)
    out = re + 1j * im
    out[0, 0, 0] = 0.0
    return out


def psi_z_spectrum(delta_k, box):
    """i k_z delta(k) / k^2 in physical units (length), k_z = n/2 plane and k = 0 zeroed."""
    n = delta_k.shape[0]
    h = n // 2
    kxx, kyy, kzz = mode_grid(n)
    sx = np.where(kxx > h, kxx - n, kxx)
    sy = np.where(kyy > h, kyy - n, kyy)
    m = (sx * sx + sy * sy + kzz * kzz).astype(F64)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where((m > 0) & (kzz < h), kzz * (box / (2.0 * np.pi)) / m, 0.0)
    return 1j * f * delta_k


def irfftn(delta_k):
    n = delta_k.shape[0]
    return np.fft.irfftn(delta_k, s=(n, n, n), axes=(0, 1, 2))


# ------------------------------------------------------------------------------------
# Poisson sampling and positions
# ------------------------------------------------------------------------------------
def _ptrs_consts(lam):
    slam = np.sqrt(lam)
    loglam = np.log(lam)
    b = 0.931 + 2.53 * slam
    a = -0.059 + 0.02483 * b
    invalpha = 1.1239 + 1.1328 / (b - 3.4)
    vr = 0.9277 - 3.6224 / (b - 2.0)
    return loglam, a, b, invalpha, vr


_LOG_FACT_TAB = [float.fromhex(h) for h in (
    "0x0.0p+0",
    "0x0.0p+0",
    "0x1.62e42fefa39ecp-1",
    "0x1.cab0bfa2a2004p+0",
    "0x1.96ca77c922cf7p+1",
    "0x1.326643c4479cap+2",
    "0x1.a51273acf01cbp+2",
    "0x1.10ce1f32dcc30p+3",
    "0x1.5358e82fcb70cp+3",
    "0x1.99a8921a7f7cep+3",
    "0x1.e357590954d14p+3",
    "0x1.180973f3a8d74p+4",
    "0x1.3fcba16d50143p+4",
    "0x1.68d5a9c3b32cdp+4",
    "0x1.930f3df162a43p+4",
    "0x1.be636a63fd346p+4")]


def log_factorial(k: float) -> float:
    """log(k!): exact table below 16, Stirling's series with four correction terms above."""
    if k < 16.0:
        return _LOG_FACT_TAB[int(k)]
    x = k + 1.0
    r = 1.0 / x
    r2 = r * r
    corr = r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 * (1.0 / 1680.0))))
    return (x - 0.5) * float(np.log(x)) - x + 0.91893853320467274178 + corr


def log_factorial_vec(k):
    """log_factorial over an array of non-negative integer-valued floats."""
    k = np.asarray(k, dtype=F64)
    tab = np.array(_LOG_FACT_TAB)
    x = k + 1.0
    r = 1.0 / x
    r2 = r * r
    corr = r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 * (1.0 / 1680.0))))
    st = (x - 0.5) * np.log(x) - x + 0.91893853320467274178 + corr
    return np.where(k < 16.0, tab[np.clip(k, 0, 15).astype(np.int64)], st)


def poisson_uniform(cells, seed):
    """Inversion uniform of cell c: block (c >> 1, 0, TAG_POISSON), words (x, y) for even c,
    (z, w) for odd c."""
    cells = np.asarray(cells, dtype=U64)
    w0, w1, w2, w3 = block(cells >> U64(1), 0, TAG_POISSON, seed)
    odd = (cells & U64(1)) == U64(1)
    return np.where(odd, u52(w2, w3), u52(w0, w1))


def poisson(lam, cells, seed):
    """Poisson draws for rates lam (FP64) of the given global cell indices."""
    lam = np.minimum(np.asarray(lam, dtype=F64), POISSON_LAM_MAX)
    cells = np.asarray(cells, dtype=U64)
    out = np.zeros(lam.shape, dtype=np.int64)
    ok = lam > 0.0
    small = ok & (lam < POISSON_INV_MAX)
    if small.any():
        ls = lam[small]
        u = poisson_uniform(cells[small], seed)
        p = np.exp(-ls)
        F = p.copy()
        k = np.zeros(ls.shape, dtype=np.int64)
        act = u > F
        while act.any():
            k = np.where(act, k + 1, k)
            p = np.where(act, p * ls / np.maximum(k, 1), p)
            F = np.where(act, F + p, F)
            act = act & (u > F) & (k < INV_KMAX)
        out[small] = k
    big = np.nonzero(ok & ~small)[0]
    if big.size:
        L = lam[big]
        cb = cells[big]
        loglam, a, b, invalpha, vr = _ptrs_consts(L)
        res = np.floor(L)                     # fallback after PTRS_MAX_IT rounds
        todo = np.arange(big.size)
        for j in range(PTRS_MAX_IT):
            if todo.size == 0:
                break
            w0, w1, w2, w3 = block(cb[todo], j, TAG_PTRS, seed)
            U = u52(w0, w1) - 0.5
            V = u52(w2, w3)
            us = 0.5 - np.abs(U)
            Lt, at, bt = L[todo], a[todo], b[todo]
            kk = np.floor((2.0 * at / us + bt) * U + Lt + 0.43)
            acc1 = (us >= 0.07) & (V <= vr[todo])
            rej = (kk < 0) | ((us < 0.013) & (V > us))
            with np.errstate(invalid="ignore", divide="ignore"):
                lhs = np.log(V) + np.log(invalpha[todo]) - np.log(at / (us * us) + bt)
                rhs = -Lt + kk * loglam[todo] - log_factorial_vec(np.maximum(kk, 0.0))
            acc = acc1 | (~rej & (lhs <= rhs))
            res[todo[acc]] = kk[acc]
            todo = todo[~acc]
        out[big] = res.astype(np.int64)
    return out


def displacements(first, count, seed):
    """Triangular in-cell offsets (src/gauss_field.py:83-88) in units of the cell size for
    particles first .. first+count-1, shape (count, 3)."""
    p = np.arange(first, first + count, dtype=U64)
    w0, w1, w2, _ = block(p, 0, TAG_DISP, seed)
    R = np.stack([2.0 * uf32(w) - 1.0 for w in (w0, w1, w2)], axis=1)
    return np.sign(R) * (1.0 - np.sqrt(np.abs(R)))


def rates(field, n, box, density, bias=0.0):
    """Per-cell Poisson rates: rho (cell volume * density) / mean(rho), FP64."""
    f = np.asarray(field, dtype=F32).reshape(-1).astype(F64)
    rho = np.exp(float(bias) * f) if bias != 0.0 else f
    mean_obj = (float(box) / n) ** 3 * float(density)
    return rho * (mean_obj / (rho.sum() / rho.size))


def populate(field, n, box, density, seed, bias=0.0, psi_z=None, growth_rate=0.0):
    """Particles (N_p, 3) float32 and per-cell counts of the populate step."""
    lam = rates(field, n, box, density, bias)
    cells = np.arange(lam.size, dtype=U64)
    counts = poisson(lam, cells, seed)
    total = int(counts.sum())
    cid = np.repeat(np.arange(lam.size, dtype=np.int64), counts)
    ix, iy, iz = np.unravel_index(cid, (n, n, n))
    d = displacements(0, total, seed)
    H = float(box) / n
    pos = np.stack([(ix + 0.5 + d[:, 0]) * H, (iy + 0.5 + d[:, 1]) * H, (iz + 0.5 + d[:, 2]) * H], axis=1)
    if psi_z is not None and growth_rate != 0.0:
        pos[:, 2] += float(growth_rate) * np.asarray(psi_z, dtype=F32).reshape(-1)[cid].astype(F64)
    return wrap_box(pos, box), counts


def wrap_box(r, box):
    """Periodic wrap into [0, L) in FP64 (values inside are kept as they are), rounded to
    float32; a value that rounds up to L becomes 0."""
    r = np.asarray(r, dtype=F64)
    inside = (r >= 0.0) & (r < box)
    with np.errstate(invalid="ignore"):
        u = r - box * np.floor(r / box)
    u = np.where(u < 0.0, u + box, u)
    u = np.where(u >= box, u - box, u)
    out = np.where(inside, r, u).astype(F32)
    out[out >= F32(box)] = 0.0
    return out
