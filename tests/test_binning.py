"""Host-side integer binning tables (jax_powspec_amd.binning) against the oracle's
jnp.histogram rule, for every m on the half-spectrum lattice (CPU only)."""
import math

import numpy as np
import pytest

from jax_powspec_amd import binning
from oracle import ref_fp64 as R


def device_rule(m, thr, jstart):
    """Python statement of the kernel's find_bin (pk_fft.hip): isqrt start + forward scan."""
    nb = thr.size - 1
    j = math.isqrt(int(m))
    b = int(jstart[j])
    while b + 1 < nb and thr[b + 1] <= m:
        b += 1
    if b < 0 or m >= thr[nb]:
        return -1
    return b


EDGE_CASES = {
    "kf_integer": lambda n: np.arange(1, n // 2 + 1, dtype=float),
    "kf_half": lambda n: 0.5 + np.arange(0, n // 2 + 1),
    "with_zero": lambda n: np.arange(0, n // 2 + 1, dtype=float),
    "narrow": lambda n: np.arange(0.3, n / 2, 0.08),            # many empty bins, > 256 bins
    "beyond_kmax": lambda n: np.array([2.0, 5.0, 1e3, 1e4]),
    "negative_first": lambda n: np.array([-3.0, 1.0, 2.5, n / 2]),
    "duplicates": lambda n: np.array([1.0, 2.0, 2.0, 3.0, 5.0, 5.0]),
    "last_edge_square": lambda n: np.array([0.5, 1.5, 3.0]),      # m = 9 sits on the closed last edge
    "sqrt2": lambda n: np.array([1.0, np.sqrt(2.0), 2.0, np.sqrt(8.0)]),
    "single_bin": lambda n: np.array([0.0, 1e9]),
}


@pytest.mark.parametrize("n", [8, 12, 24, 64])
@pytest.mark.parametrize("case", sorted(EDGE_CASES))
def test_thresholds_equal_histogram_on_every_m(n, case):
    edges = EDGE_CASES[case](n)
    thr, jst = binning.thresholds_from_grid_edges(edges, n)
    m = np.arange(binning.max_m(n) + 1)
    want = R.jnp_histogram_slots(np.sqrt(m.astype(np.float64)), edges)
    np.testing.assert_array_equal(binning.slots_for_m(m, thr), want)
    # the kernel's isqrt-start search gives the same slots
    got = np.array([device_rule(v, thr, jst) for v in m[:: max(1, m.size // 4000)]])
    np.testing.assert_array_equal(got, want[:: max(1, m.size // 4000)])


def test_random_edges_against_histogram():
    rng = np.random.default_rng(1)
    n = 48
    m = np.arange(binning.max_m(n) + 1)
    for _ in range(30):
        edges = np.sort(rng.uniform(-2, 45, size=rng.integers(2, 40)))
        if rng.random() < 0.5:
            edges = np.round(edges * 2) / 2          # ties on half-integers / integers
        thr, jst = binning.thresholds_from_grid_edges(edges, n)
        want = R.jnp_histogram_slots(np.sqrt(m.astype(np.float64)), edges)
        np.testing.assert_array_equal(binning.slots_for_m(m, thr), want)


def test_caller_edge_patterns_in_physical_units():
    # edge constructions of the reference's callers (tests/create_lognormal.py:68,
    # tests/lognormal.py:70, tests/correlations.py:76) with k_edges / kF in FP64
    for box, n, k_edges in [
        (1000.0, 128, np.arange(2 * np.pi / 1000.0, np.pi * 128 / 1000.0, 2 * np.pi / 1000.0)),
        (1000.0, 64, np.arange(0.003, np.pi * 64 / 1000.0, 0.0005)),
        (1500.0, 32, np.arange(1e-4, 5, 0.2e-2)),
    ]:
        e = binning.grid_edges(k_edges, box)
        thr, _ = binning.thresholds_from_grid_edges(e, n)
        m = np.arange(binning.max_m(n) + 1)
        want = R.jnp_histogram_slots(np.sqrt(m.astype(np.float64)), e)
        np.testing.assert_array_equal(binning.slots_for_m(m, thr), want)


def test_fundamental_edges_give_isqrt_bins():
    n = 64
    thr, _ = binning.thresholds_from_grid_edges(binning.fundamental_grid_edges(n), n)
    m = np.arange(binning.max_m(n) + 1)
    want = np.array([math.isqrt(int(v)) - 1 for v in m])
    np.testing.assert_array_equal(binning.slots_for_m(m, thr), want)
    assert thr.size - 1 == math.isqrt(3 * (n // 2) ** 2)


def test_decreasing_edges_rejected():
    with pytest.raises(ValueError):
        binning.thresholds_from_grid_edges(np.array([1.0, 3.0, 2.0]), 16)
    with pytest.raises(ValueError):
        binning.thresholds_from_grid_edges(np.array([1.0]), 16)


@pytest.mark.parametrize("n", [8, 12, 24, 64])
@pytest.mark.parametrize("case", sorted(EDGE_CASES))
def test_mode_tables_equal_brute_force_counts(n, case):
    """Exact per-bin mode counts and sum of |k| from the m-multiplicity table equal a
    brute-force pass over every stored half-spectrum entry under jnp.histogram's rule."""
    edges = EDGE_CASES[case](n)
    thr, _ = binning.thresholds_from_grid_edges(edges, n)
    modes, ksum = binning.bin_modes(thr, n)
    kx, ky, kz = R.half_spectrum_grids(n)
    m = np.broadcast_to(kx * kx + ky * ky + kz * kz, (n, n, n // 2 + 1)).reshape(-1)
    k = np.sqrt(m.astype(np.float64))
    slot = R.jnp_histogram_slots(k, edges)
    nb = edges.size - 1
    sel = slot >= 0
    want_c = np.bincount(slot[sel], minlength=nb)[:nb]
    want_k = np.bincount(slot[sel], weights=k[sel], minlength=nb)[:nb]
    assert modes.dtype == np.int64 and modes.shape == (nb,)
    np.testing.assert_array_equal(modes, want_c)
    np.testing.assert_allclose(ksum, want_k, rtol=1e-13, atol=0)
    assert modes.sum() == sel.sum()


def test_mode_multiplicity_totals():
    # every stored entry counted once: n * n * (n/2 + 1) in total; m = 0 only at k = 0
    for n in (8, 10, 64, 512):
        mult = binning.mode_multiplicity(n)
        assert mult.sum() == n * n * (n // 2 + 1)
        assert mult[0] == 1 and mult[1] == 5     # k = 0; the six unit vectors less kz = -1
