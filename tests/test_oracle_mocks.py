"""Pins the mock-generator oracle (oracle/mocks.py) with checks that do not restate it.

* Philox4x32-10 against the generator's published known-answer vectors (Random123).
* P(k) interpolation against the reference's own bisection loop
  (src/gauss_field.py:39-46), written as that scalar loop.
* The vectorised half-spectrum fill against a scalar emulation of the reference's nested
  loop and its "fill if still zero" Hermitian rule (src/gauss_field.py:24-78); every
  mode there draws from its own loop index, and the first of (k, -k) keeps its draws.
* Hermitian symmetry through NumPy's FFT round trip; |delta_k|^2 = P (n^2/L)^3 without
  Rayleigh sampling; E[-ln u] = 1 with it.
* Poisson inversion against scipy's quantile function; both regimes against the Poisson
  pmf (chi-square); the displacement law against the triangular CDF (Kolmogorov-Smirnov).
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import stats

from oracle import mocks as M

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    got = tuple(int(v) for v in M.philox4x32(*ctr, *key))
    assert got == want


def test_uniform_maps_stay_open():
    w = np.array([0, 1, 255, 256, 0xFFFFFFFF], dtype=np.uint64)
    u = M.uf32(w)
    assert np.all(u > 0) and np.all(u < 1)
    assert np.all(u.astype(np.float32).astype(np.float64) == u)       # float32-exact
    v = M.u52(w, w)
    assert np.all(v > 0) and np.all(v < 1)
    assert M.u52(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53


def _bisect_interp(kmod, kf, Pkf):
    # the reference's scalar loop (src/gauss_field.py:39-46)
    lmin, lmax = 0, len(kf) - 1
    while lmax - lmin > 1:
        l = (lmin + lmax) // 2
        if kf[l] < kmod:
            lmin = l
        else:
            lmax = l
    return (Pkf[lmax] - Pkf[lmin]) / (kf[lmax] - kf[lmin]) * (kmod - kf[lmin]) + Pkf[lmin]


def test_pk_interp_matches_bisection():
    rng = np.random.default_rng(3)
    kf = np.sort(rng.random(37)) * 0.8 + 0.01
    kf = np.unique(kf)
    Pkf = rng.random(kf.size) * 1e4
    ks = np.concatenate([rng.random(400) * 1.0, kf, [0.0, 1e-9, 5.0]])
    got = M.pk_interp(ks, kf, Pkf)
    want = np.array([_bisect_interp(k, kf, Pkf) for k in ks])
    assert np.array_equal(got, want)


def _reference_loop_fill(n, kf, Pkf, rayleigh, seed, box):
    """Scalar emulation of src/gauss_field.py:24-78 with draws from the mode's own loop index."""
    h = n // 2
    d = np.zeros((n, n, h + 1), dtype=np.complex128)
    amp_tab = M.amplitude_table(n, box, kf, Pkf)
    for kxx in range(n):
        kx = kxx - n if kxx > h else kxx
        kxx_m = n - kx if kx > 0 else -kx
        for kyy in range(n):
            ky = kyy - n if kyy > h else kyy
            kyy_m = n - ky if ky > 0 else -ky
            for kzz in range(h + 1):
                kz = kzz
                idx = (kxx * n + kyy) * (h + 1) + kzz
                w0, w1, _, _ = M.block(idx, 0, M.TAG_FIELD, seed)
                phase = 2.0 * math.pi * float(M.uf32(w0))
                amplitude = math.sqrt(-math.log(float(M.uf32(w1)))) if rayleigh == 1 else 1.0
                amplitude *= amp_tab[kx * kx + ky * ky + kz * kz]
                re, im = amplitude * math.cos(phase), amplitude * math.sin(phase)
                if d[kxx, kyy, kzz] == 0:
                    d[kxx, kyy, kzz] = re + 1j * im
                    if kz == 0 or kz == h:
                        if d[kxx_m, kyy_m, kzz] == 0:
                            d[kxx_m, kyy_m, kzz] = re - 1j * im
                        if kxx_m == kxx and kyy_m == kyy:
                            d[kxx, kyy, kzz] = amplitude + 0j
    d[0, 0, 0] = 0
    return d


def _table(box):
    kf = np.linspace(1e-4, 2.0, 300)
    Pkf = 2e4 * (kf / 0.05) / (1.0 + (kf / 0.05) ** 2.5)
    return kf, Pkf


@pytest.mark.parametrize("n,rayleigh,seed", [(8, 0, 11), (8, 1, 12), (12, 1, 2**40 + 5), (16, 1, 7)])
def test_gaussian_field_matches_reference_loop(n, rayleigh, seed):
    box = 300.0
    kf, Pkf = _table(box)
    got = M.gaussian_field(n, kf, Pkf, rayleigh, seed, box)
    want = _reference_loop_fill(n, kf, Pkf, rayleigh, seed, box)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("n", [16, 24, 32])
def test_gaussian_field_is_hermitian(n):
    box = 500.0
    kf, Pkf = _table(box)
    d = M.gaussian_field(n, kf, Pkf, 1, 99, box)
    g = np.fft.irfftn(d, s=(n, n, n), axes=(0, 1, 2))
    back = np.fft.rfftn(g)
    np.testing.assert_allclose(back, d, rtol=0, atol=1e-10 * np.abs(d).max())
    assert d[0, 0, 0] == 0


def test_gaussian_field_amplitudes():
    n, box = 32, 800.0
    kf, Pkf = _table(box)
    d = M.gaussian_field(n, kf, Pkf, 0, 5, box)
    kxx, kyy, kzz = np.broadcast_arrays(*M.mode_grid(n))
    h = n // 2
    sx = np.where(kxx > h, kxx - n, kxx)
    sy = np.where(kyy > h, kyy - n, kyy)
    kmod = np.sqrt(sx**2 + sy**2 + kzz**2) * 2 * np.pi / box
    P = np.interp(kmod, kf, Pkf) * (n**2 / box) ** 3      # kf covers every |k| here
    P[0, 0, 0] = 0.0
    np.testing.assert_allclose(np.abs(d) ** 2, P, rtol=1e-10, atol=0)
    dr = M.gaussian_field(n, kf, Pkf, 1, 6, box)
    ratio = (np.abs(dr) ** 2)[P > 0] / P[P > 0]
    assert abs(ratio.mean() - 1.0) < 5 * ratio.std() / math.sqrt(ratio.size)


def test_psi_z_is_minus_gradient_inverse():
    # div psi = -delta for a field without a Nyquist-z component: i k . psi(k) = -delta(k)
    n, box = 16, 200.0
    kf, Pkf = _table(box)
    d = M.gaussian_field(n, kf, Pkf, 1, 3, box)
    d[:, :, n // 2] = 0.0
    kxx, kyy, kzz = M.mode_grid(n)
    psi = M.psi_z_spectrum(d, box)
    kz = kzz * 2 * np.pi / box
    h = n // 2
    sx = np.where(kxx > h, kxx - n, kxx) * 2 * np.pi / box
    sy = np.where(kyy > h, kyy - n, kyy) * 2 * np.pi / box
    k2 = sx**2 + sy**2 + kz**2
    # psi_x, psi_y analogues from the same rule: i k_a delta / k^2
    with np.errstate(divide="ignore", invalid="ignore"):
        px = np.where(k2 > 0, 1j * sx * d / k2, 0)
        py = np.where(k2 > 0, 1j * sy * d / k2, 0)
    div = 1j * (sx * px + sy * py + kz * psi)
    np.testing.assert_allclose(div, -d, rtol=0, atol=1e-10 * np.abs(d).max())


def test_poisson_inversion_is_the_quantile():
    rng = np.random.default_rng(1)
    lam = rng.random(20000) * 9.99
    cells = np.arange(lam.size, dtype=np.uint64) * 7 + 3
    k = M.poisson(lam, cells, 42)
    u = M.poisson_uniform(cells, 42)
    q = stats.poisson.ppf(u, lam).astype(np.int64)
    # the inversion sums the pmf forward in FP64; scipy's quantile differs only where u is
    # within rounding of a CDF value
    bad = np.nonzero(k != q)[0]
    assert bad.size <= 2, (bad, lam[bad], u[bad], k[bad], q[bad])


@pytest.mark.parametrize("lam", [0.15, 2.5, 9.5, 10.0, 31.0, 400.0])
def test_poisson_distribution(lam):
    ncell = 40000 if lam < 10 else 6000
    k = M.poisson(np.full(ncell, lam), np.arange(ncell, dtype=np.uint64) + 1000, 2024)
    lo, hi = stats.poisson.ppf([1e-4, 1 - 1e-4], lam).astype(int)
    edges = np.arange(lo, hi + 2)
    obs = np.array([np.sum(k < lo)] + [np.sum(k == v) for v in edges[:-1]] + [np.sum(k > hi)])
    p = np.concatenate([[stats.poisson.cdf(lo - 1, lam)], stats.poisson.pmf(edges[:-1], lam),
                        [stats.poisson.sf(hi, lam)]])
    exp = p * ncell
    # pool sparse classes
    keep = exp >= 5
    o = np.concatenate([obs[keep], [obs[~keep].sum()]])
    e = np.concatenate([exp[keep], [exp[~keep].sum()]])
    if e[-1] < 5:
        o, e = o[:-1], e[:-1]
    chi2 = ((o - e) ** 2 / e).sum()
    pval = stats.chi2.sf(chi2, len(o) - 1)
    assert pval > 1e-4, (lam, chi2, len(o), pval)


def test_displacement_law_is_triangular():
    d = M.displacements(0, 60000, 77)
    assert np.all(np.abs(d) < 1)
    cdf = lambda x: np.where(x < 0, 0.5 * (1 + x) ** 2, 1 - 0.5 * (1 - x) ** 2)
    for a in range(3):
        assert stats.kstest(d[:, a], cdf).pvalue > 1e-4


def test_populate_counts_and_cells():
    n, box, density = 16, 100.0, 0.05
    rng = np.random.default_rng(0)
    g = rng.standard_normal((n, n, n)).astype(np.float32)
    pos, counts = M.populate(g, n, box, density, seed=9, bias=1.2)
    lam = M.rates(g, n, box, density, 1.2)
    assert abs(lam.sum() - density * box**3) < 1e-9 * lam.sum()
    assert pos.shape == (counts.sum(), 3) and pos.dtype == np.float32
    assert np.all(pos >= 0) and np.all(pos < box)
    # each particle lies within one cell of its cell's centre (periodic)
    cid = np.repeat(np.arange(n**3), counts)
    centre = (np.stack(np.unravel_index(cid, (n, n, n)), axis=1) + 0.5) * (box / n)
    dd = (pos - centre + box / 2) % box - box / 2
    assert np.all(np.abs(dd) < box / n * (1 + 1e-5))
    tot = counts.sum()
    assert abs(tot - lam.sum()) < 6 * math.sqrt(lam.sum())


def test_log_factorial():
    for k in list(range(0, 200)) + [500, 1000, 12345, 10**6]:
        want = math.lgamma(k + 1.0)
        got = M.log_factorial(float(k))
        assert abs(got - want) <= 2e-15 * max(1.0, abs(want)), (k, got, want)


def test_wrap_box_stays_inside():
    box = 1000.0
    r = np.array([0.0, -0.0, 1e-300, -1e-300, box, box - 1e-13, -box, 2 * box - 1e-9, -3.5 * box, 123.25,
                  np.nextafter(box, 0), np.nextafter(0.0, -1)])
    w = M.wrap_box(r, box)
    assert w.dtype == np.float32
    assert np.all(w >= 0) and np.all(w < np.float32(box))
    assert w[9] == np.float32(123.25)
    # periodic images agree up to float32 rounding of the box
    d = np.abs(w.astype(np.float64) - (r % box))
    assert np.all(np.minimum(d, box - d) <= 1e-4)


def test_log_factorial_vec_matches_scalar():
    ks = np.array([0, 1, 5, 15, 16, 17, 40, 1000, 123456], dtype=np.float64)
    v = M.log_factorial_vec(ks)
    for k, x in zip(ks, v):
        assert abs(x - M.log_factorial(float(k))) <= 1e-15 * max(1.0, abs(x))
