"""Pins the configuration-space oracle (oracle/ref_fp64.py: xi_grid, xi_vec,
xi_vec_fundamental) and the full-grid mode tables with checks that do not restate them:

* xi_grid without the window is the brute-force periodic autocorrelation
  sum_x delta(x) delta(x + s);
* bin populations equal numpy.histogram of |s| over the full grid (an independent
  implementation of the same rule) and the enumeration of the full grid;
* the sum rule sum_s xi_grid(s) = |delta_k(0)|^2 = (sum delta)^2 (edges covering every s);
* the fundamental variant's shells are int(|s|) of an explicit enumeration, its r the
  cell-averaged |s|; the shell holding s = 0 gets NaN l = 2, 4 when an edge admits it.
"""
import numpy as np
import pytest

from jax_powspec_amd import binning
from oracle import ref_fp64 as R


def _signed(n):
    r = np.arange(n)
    return np.where(r > n // 2, r - n, r)


@pytest.mark.parametrize("n", [6, 8])
def test_xi_grid_is_the_autocorrelation(n):
    rng = np.random.default_rng(n)
    d = rng.standard_normal((n, n, n))
    g = R.xi_grid(d, p=0)
    bf = np.zeros_like(g)
    for a in range(n):
        for b in range(n):
            for c in range(n):
                bf[a, b, c] = (d * np.roll(d, (-a, -b, -c), axis=(0, 1, 2))).sum()
    np.testing.assert_allclose(g, bf, rtol=0, atol=1e-10 * np.abs(bf).max())


@pytest.mark.parametrize("n,box,edges", [(8, 100.0, np.arange(1e-3, 60, 7.0)), (12, 300.0, np.linspace(0, 150, 9)),
                                         (16, 50.0, np.array([0.5, 3.125, 9.375, 40.0]))])
def test_xi_bin_populations(n, box, edges):
    d = np.random.default_rng(1).standard_normal((n, n, n))
    _, _, N, ex = R.xi_vec(d, box, edges)
    s = _signed(n)
    k = np.sqrt((s[:, None, None] ** 2 + s[None, :, None] ** 2 + s[None, None, :] ** 2).astype(np.float64))
    ke = edges * n / box
    want, _ = np.histogram(k.reshape(-1), bins=ke)
    np.testing.assert_array_equal(ex["counts"], want)
    thr, _ = binning.thresholds_from_grid_edges(ke, n)
    modes, _ = binning.bin_modes(thr, n, full=True)
    np.testing.assert_array_equal(modes, want)
    assert np.all(np.isinf(N[want == 0]))


@pytest.mark.parametrize("n", [8, 16])
def test_xi_sum_rule(n):
    d = 1.0 + 0.3 * np.random.default_rng(2).standard_normal((n, n, n))
    box = 100.0
    edges = np.array([-1.0, 2.0 * box])     # every separation
    _, xi, N, _ = R.xi_vec(d, box, edges, p=0)
    total = xi[0, 0] * N[0] * n ** 3
    assert abs(total - d.sum() ** 2) <= 1e-10 * d.sum() ** 2
    assert np.isnan(xi[0, 1]) and np.isnan(xi[0, 2])      # the s = 0 cell is in the bin


@pytest.mark.parametrize("n", [8, 12, 16])
def test_xi_fundamental_shells(n):
    d = np.random.default_rng(3).standard_normal((n, n, n))
    box = 200.0
    r, xi, N, ex = R.xi_vec_fundamental(d, box)
    s = _signed(n)
    m = (s[:, None, None] ** 2 + s[None, :, None] ** 2 + s[None, None, :] ** 2).reshape(-1)
    j = np.array([int(np.sqrt(v)) for v in m])
    kmax = int(np.sqrt(3 * (n // 2) ** 2))
    for sh in range(1, kmax + 1):
        sel = j == sh
        assert ex["counts"][sh - 1] == sel.sum()
        if sel.any():
            assert abs(r[sh - 1] - np.sqrt(m[sel]).mean() * box / n) < 1e-9 * box
    thr, _ = binning.thresholds_from_grid_edges(binning.fundamental_grid_edges(n), n)
    modes, _ = binning.bin_modes(thr, n, full=True)
    np.testing.assert_array_equal(modes, ex["counts"])


@pytest.mark.parametrize("n", [6, 8, 12])
def test_full_grid_multiplicity(n):
    s = _signed(n)
    m = (s[:, None, None] ** 2 + s[None, :, None] ** 2 + s[None, None, :] ** 2).reshape(-1)
    np.testing.assert_array_equal(np.bincount(m, minlength=binning.max_m(n) + 1), binning.full_grid_multiplicity(n))
