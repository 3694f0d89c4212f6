"""Pins the bispectrum oracle (oracle/ref_fp64.py: bispec) with checks that do not
restate it:

* Parseval: sum_x delta_i(x)^2 = (1/n^3) sum over the full spectrum of |shell delta_k W|^2
  (so P_i is the shell average of the pinned power spectrum);
* the triple sum against a brute-force enumeration of closed triangles
  sum_x d1 d2 d3 = n^-6 sum_{q1 + q2 + q3 = 0} A1(q1) A2(q2) A3(q3) on the full spectrum;
* the shell bounds (k_i/kF -/+ 1) and integer thresholds against a direct comparison.
"""
import numpy as np
import pytest

from jax_powspec_amd import correlations as CO
from oracle import ref_fp64 as R


def full_spectrum(n):
    ki = np.where(np.arange(n) > n // 2, np.arange(n) - n, np.arange(n))
    return ki[:, None, None], ki[None, :, None], ki[None, None, :]


@pytest.mark.parametrize("n", [8, 12])
def test_bispec_parseval_and_triangles(n):
    rng = np.random.default_rng(n)
    d = rng.standard_normal((n, n, n))
    L = 100.0
    kF = 2 * np.pi / L
    theta = np.array([0.3, 1.4, 2.9])
    k_all, Pk, _, B, Q, ex = R.bispec(d, L, 1.5 * kF, 2.0 * kF, theta, p=0)
    D = np.fft.fftn(d)
    kx, ky, kz = full_spectrum(n)
    k = np.sqrt((kx ** 2 + ky ** 2 + kz ** 2).astype(np.float64))
    _, kmin, kmax = R.bispec_shells(L, 1.5 * kF, 2.0 * kF, theta)
    A = [np.where((k >= lo) & (k < hi), D, 0.0) for lo, hi in zip(kmin, kmax)]
    M = [((k >= lo) & (k < hi)).astype(np.float64) for lo, hi in zip(kmin, kmax)]
    for i in range(len(A)):
        assert abs((np.abs(A[i]) ** 2).sum() / n ** 3 - ex["sum_d2"][i]) <= 1e-9 * ex["sum_d2"][i]
        assert abs((M[i] ** 2).sum() / n ** 3 - ex["sum_I2"][i]) <= 1e-9 * max(ex["sum_I2"][i], 1e-30)
    # closed triangles q1 + q2 + q3 = 0 (mod n), by brute force over q1, q2
    q = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    for i in range(2, len(A)):
        a1, a2, a3 = (X.reshape(-1) for X in (A[0], A[1], A[i]))
        m1, m2, m3 = (X.reshape(-1) for X in (M[0], M[1], M[i]))
        t = 0.0
        c = 0.0
        nz1 = np.nonzero(m1)[0]
        nz2 = np.nonzero(m2)[0]
        for i1 in nz1:
            q3 = (-(q[i1][None, :] + q[nz2])) % n
            j3 = (q3[:, 0] * n + q3[:, 1]) * n + q3[:, 2]
            t += (a1[i1] * a2[nz2] * a3[j3]).sum()
            c += (m1[i1] * m2[nz2] * m3[j3]).sum()
        assert abs(t.real / n ** 6 - ex["triples"][i - 2]) <= 1e-9 * (abs(ex["triples"][i - 2]) + 1e-9)
        assert abs(c / n ** 6 - ex["triangles"][i - 2]) <= 1e-12 + 1e-9 * abs(ex["triangles"][i - 2])


def test_bispec_shell_thresholds():
    n, L = 32, 300.0
    kF = 2 * np.pi / L
    theta = np.linspace(0, np.pi, 7)
    k_all, th, lo, hi = CO.bispec_shells(n, L, 3 * kF, 5 * kF, theta)
    m = np.arange(3 * (n // 2) ** 2 + 1)
    k = np.sqrt(m.astype(np.float64))
    for i, ka in enumerate(k_all):
        inside = (k >= ka / kF - 1) & (k < ka / kF + 1)
        got = (m >= lo[i]) & (m < hi[i])
        np.testing.assert_array_equal(got, inside)


def test_bispec_window_is_the_reference_cic_correction():
    # the reference multiplies delta_k by (1 / sinc(k / n))^2 per axis (src/correlations.py:362-373):
    # with p = 1, P_i is the shell average of |delta_k|^2 times that factor squared
    n, L = 8, 100.0
    d = np.random.default_rng(4).standard_normal((n, n, n))
    kF = 2 * np.pi / L
    _, Pk, _, _, _, ex = R.bispec(d, L, 1.5 * kF, 2.0 * kF, np.array([1.0]), p=1)
    D = np.fft.fftn(d)
    kx, ky, kz = full_spectrum(n)
    W = ((1 / np.sinc(kx / n)) * (1 / np.sinc(ky / n)) * (1 / np.sinc(kz / n))) ** 2
    k = np.sqrt((kx ** 2 + ky ** 2 + kz ** 2).astype(np.float64))
    _, kmin, kmax = R.bispec_shells(L, 1.5 * kF, 2.0 * kF, np.array([1.0]))
    for i, (lo, hi) in enumerate(zip(kmin, kmax)):
        sh = (k >= lo) & (k < hi)
        assert abs((np.abs(D * W) ** 2)[sh].sum() / n ** 3 - ex["sum_d2"][i]) <= 1e-9 * ex["sum_d2"][i]
