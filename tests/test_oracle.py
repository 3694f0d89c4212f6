"""Pins of the CPU oracle (oracle/ref_fp64.py) against checks that do not reuse its code.

The reference ships no golden vectors or fixtures, and jax is absent, so the oracle is
pinned by:
  * literal per-particle restatements of src/mas.py (loops + explicit branches) for the
    reference stencil in all four wrap modes, including out-of-range particles;
  * closed forms: grid-point particles, the src/mas.py:149 corner weight, mass sums,
    TSC weights;
  * a dense DFT (brute force) for the FFT, plane waves with known multipoles for the
    deconvolution / Legendre / normalisation, Parseval;
  * exact (rational) lattice enumeration of mode counts, including ties on edges;
  * the shot-noise expectation of uniform catalogues for textbook CIC and TSC
    (statistical; pins the window exponent and the (L/n^2)^3 normalisation).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import ref_fp64 as R

F32 = np.float32


# ----------------------------------------------------------------------------------------
# mass assignment
# ----------------------------------------------------------------------------------------
def _jax_norm(i, n):
    i = i + n if i < 0 else i
    return i if 0 <= i < n else None


def brute_ref(x, y, z, w, xmin, box, n, mode):
    """Per-particle restatement of src/mas.py:40-81 / :103-151 with explicit branches."""
    mesh = np.zeros((n, n, n))
    bin_size = F32(box) / F32(n)
    inv = F32(1.0) / bin_size
    for p in range(len(x)):
        pos = [(F32(c[p]) - F32(xmin)) * inv for c in (x, y, z)]
        i = [int(q) for q in pos]                   # int32(): truncation toward zero
        dd = [F32(q - F32(k)) for q, k in zip(pos, i)]
        md = [F32(F32(1.0) - d) for d in dd]
        ip = []
        for a in range(3):
            q = i[a] + 1
            if mode == R.WRAP_VEC_TRUE:
                q = (q + n) % n                     # Python % is floor mod like jnp
            elif mode == R.WRAP_VEC_FALSE:
                if q >= n:
                    q = 0
            elif mode == R.WRAP_SCAN_TRUE:
                if q >= n:
                    q = q - n
            else:
                if q >= n:
                    q = 0
                    dd[a] = F32(0.0)
            ip.append(q)
        W = float(w[p])
        md = [float(v) for v in md]           # FP64 products of the float32 factors
        dd = [float(v) for v in dd]
        corners = [
            ((i[0], i[1], i[2]), md[0] * md[1] * md[2]),
            ((ip[0], i[1], i[2]), dd[0] * md[1] * md[2]),
            ((i[0], ip[1], i[2]), md[0] * dd[1] * md[2]),
            ((i[0], i[1], ip[2]), md[0] * md[1] * dd[2]),
            ((ip[0], ip[1], i[2]), dd[0] * dd[1] * md[2]),
            ((ip[0], i[1], ip[2]), dd[0] * md[1] * dd[2]),
            ((i[0], ip[1], ip[2]), md[0] * md[1] * dd[2]),
            ((ip[0], ip[1], ip[2]), dd[0] * dd[1] * dd[2]),
        ]
        for (a, b, c), wt in corners:
            a, b, c = _jax_norm(a, n), _jax_norm(b, n), _jax_norm(c, n)
            if a is None or b is None or c is None:
                continue
            mesh[a, b, c] += float(wt) * W
    return mesh


def _edge_catalogue(rng, n_rand, box):
    pts = list(rng.random((n_rand, 3)) * box)
    H = box / 16
    specials = [0.0, box - 1e-3, box, -0.3 * H, -1.5 * H, 1.7 * box, 2 * box + 0.1, box - 0.5 * H, 0.5 * H]
    for s in specials:
        pts.append([s, rng.random() * box, rng.random() * box])
        pts.append([rng.random() * box, s, rng.random() * box])
        pts.append([rng.random() * box, rng.random() * box, s])
    pts.append([box, box, box])
    pts.append([-1.5 * H, -1.5 * H, -1.5 * H])
    return np.asarray(pts, dtype=F32)


@pytest.mark.parametrize("mode", [R.WRAP_VEC_TRUE, R.WRAP_VEC_FALSE, R.WRAP_SCAN_TRUE, R.WRAP_SCAN_FALSE])
def test_ref_stencil_matches_per_particle_restatement(mode):
    rng = np.random.default_rng(11 + mode)
    box, n = 100.0, 16
    p = _edge_catalogue(rng, 300, box)
    w = (0.5 + rng.random(len(p))).astype(F32)
    want = brute_ref(p[:, 0], p[:, 1], p[:, 2], w, 0.0, box, n, mode)
    got = R._accumulate(np.zeros((n, n, n)), n, R.cic_ref_deposits(p[:, 0], p[:, 1], p[:, 2], w, 0.0, 0.0, 0.0, box, n, mode))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_vec_and_scan_entry_points_use_their_modes():
    rng = np.random.default_rng(3)
    box, n = 50.0, 8
    p = _edge_catalogue(rng, 50, box)
    w = np.ones(len(p), F32)
    z = np.zeros((n, n, n))
    for fn, wrap, mode in ((R.cic_mas_vec, True, R.WRAP_VEC_TRUE), (R.cic_mas_vec, False, R.WRAP_VEC_FALSE),
                           (R.cic_mas, True, R.WRAP_SCAN_TRUE), (R.cic_mas, False, R.WRAP_SCAN_FALSE)):
        got = fn(z, p[:, 0], p[:, 1], p[:, 2], w, len(p), 0.0, 0.0, 0.0, box, n, wrap)
        np.testing.assert_allclose(got, brute_ref(p[:, 0], p[:, 1], p[:, 2], w, 0.0, box, n, mode), atol=1e-12)


def test_ref_corner_weight_of_line_149():
    # one particle with fractions (0.25, 0.5+0.1, 0.75) in cell (2, 3, 4), box = n: pos = x
    n = 8
    x, y, z = F32(2.25), F32(3.6), F32(4.75)
    m = R.cic_mas_vec(np.zeros((n, n, n)), [x], [y], [z], [F32(2.0)], 1, 0.0, 0.0, 0.0, float(n), n, True)
    ddx, ddy, ddz = 0.25, float(F32(3.6) - F32(3)), 0.75
    mdx, mdy = 1 - ddx, 1 - ddy
    assert m[2, 4, 5] == pytest.approx(2.0 * mdx * mdy * ddz, rel=1e-6)      # not mdx*ddy*ddz
    assert m[2, 4, 5] != pytest.approx(2.0 * mdx * ddy * ddz, rel=1e-3)
    # total mass w * (1 + mdx*ddz*(1 - 2 ddy))
    assert m.sum() == pytest.approx(2.0 * (1 + mdx * ddz * (1 - 2 * ddy)), rel=1e-6)


def test_grid_point_particles():
    n, box = 16, 16.0
    pts = np.array([[3.0, 5.0, 7.0], [0.0, 0.0, 0.0], [15.0, 1.0, 2.0]], F32)
    w = np.array([1.5, 2.0, 0.5], F32)
    for fn in (lambda: R.cic_mas_vec(np.zeros((n,) * 3), *pts.T, w, 3, 0., 0., 0., box, n, True),
               lambda: R.cic_mesh(*pts.T, w, 0, 0, 0, box, n)):
        m = fn()
        for (a, b, c), ww in zip(pts.astype(int), w):
            assert m[a, b, c] == pytest.approx(ww)
        assert m.sum() == pytest.approx(w.sum())
        assert np.count_nonzero(m) == 3
    # TSC at a grid point: 1/8, 3/4, 1/8 per axis
    m = R.tsc_mesh(*pts[:1].T, w[:1], 0, 0, 0, box, n)
    wt = {-1: 0.125, 0: 0.75, 1: 0.125}
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                assert m[3 + dx, 5 + dy, 7 + dz] == pytest.approx(1.5 * wt[dx] * wt[dy] * wt[dz])


def test_tsc_weights_closed_form_and_periodicity():
    n, box = 8, 8.0
    # d = pos - round(pos) = -0.3 at pos = 6.7 (nearest point 7), periodic neighbour 0
    m = R.tsc_mesh([F32(6.7)], [F32(1.0)], [F32(1.0)], [F32(1.0)], 0, 0, 0, box, n)
    d = float(F32(6.7)) - 7.0
    line = m[:, 1, 1] / 0.75 / 0.75
    assert line[6] == pytest.approx(0.5 * (0.5 - d) ** 2)
    assert line[7] == pytest.approx(0.75 - d * d)
    assert line[0] == pytest.approx(0.5 * (0.5 + d) ** 2)


def test_mass_conservation_cic_tsc():
    rng = np.random.default_rng(5)
    p = (rng.random((2000, 3)) * 300).astype(F32)
    w = (rng.random(2000) + 0.1).astype(F32)
    for fn in (R.cic_mesh, R.tsc_mesh):
        m = fn(*p.T, w, 0, 0, 0, 300.0, 12)
        assert m.sum() == pytest.approx(float(w.astype(np.float64).sum()), rel=1e-12)


def test_wrap_false_equals_true_inside_box():
    # src/mas.py:125-126: ddx is tested after ixp was reset, so wrap=False wraps too
    rng = np.random.default_rng(9)
    p = (rng.random((500, 3)) * 99.0).astype(F32)
    w = np.ones(500, F32)
    a = R.cic_mas_vec(np.zeros((10,) * 3), *p.T, w, 500, 0., 0., 0., 100., 10, True)
    b = R.cic_mas_vec(np.zeros((10,) * 3), *p.T, w, 500, 0., 0., 0., 100., 10, False)
    np.testing.assert_array_equal(a, b)


def test_scan_nowrap_last_cell():
    # src/mas.py:30-38: in the last cell the +1 corners get ddx = 0, base keeps mdx = 1 - ddx
    n = 8
    m = R.cic_mas(np.zeros((n,) * 3), [F32(7.25)], [F32(2.0)], [F32(3.0)], [F32(1.0)], 1, 0., 0., 0., 8., n, False)
    assert m[7, 2, 3] == pytest.approx(0.75)
    assert m.sum() == pytest.approx(0.75)


# ----------------------------------------------------------------------------------------
# FFT / power spectrum
# ----------------------------------------------------------------------------------------
def dense_rdft(a):
    n = a.shape[0]
    idx = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(idx, idx) / n)
    full = np.einsum("ai,bj,ck,ijk->abc", F, F, F, a.astype(np.complex128))
    return full[:, :, : n // 2 + 1]


@pytest.mark.parametrize("n", [6, 8])
def test_spectrum_matches_dense_dft(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n))
    d2, k, kz2, m = R.spectrum_power(a, p=0)
    want = np.abs(dense_rdft(a)) ** 2
    np.testing.assert_allclose(d2, want.reshape(-1), rtol=1e-10, atol=1e-9)


def test_parseval_weights():
    rng = np.random.default_rng(1)
    n = 12
    a = rng.standard_normal((n, n, n))
    d2, k, kz2, m = R.spectrum_power(a, p=0)
    d2 = d2.reshape(n, n, n // 2 + 1)
    full = d2[:, :, 0].sum() + d2[:, :, n // 2].sum() + 2 * d2[:, :, 1:n // 2].sum()
    assert full == pytest.approx(n ** 3 * (a * a).sum(), rel=1e-12)


def _lattice_counts_exact(n, edges, rounded_sqrt=False):
    """Stored half-spectrum modes per bin, by enumeration.

    rounded_sqrt=False compares sqrt(m) with the edges exactly (rationals); True compares
    the correctly rounded math.sqrt(m), which is what a float histogram does and differs
    only when an edge lies within half an ulp of some sqrt(m).
    """
    ki = [i if i <= n // 2 else i - n for i in range(n)]
    E = [Fraction(float(e)) for e in edges]
    E2 = [e * e if e > 0 else None for e in E]
    nb = len(edges) - 1
    counts = np.zeros(nb, np.int64)
    from collections import Counter
    ms = Counter(kx * kx + ky * ky + kz * kz for kx in ki for ky in ki for kz in range(n // 2 + 1))
    for m, c in ms.items():
        if rounded_sqrt:
            r = Fraction(math.sqrt(m))
            ge = [r >= E[i] for i in range(nb + 1)]
            eq_last = r == E[-1]
        else:
            # sqrt(m) >= e  <=>  e <= 0 or m >= e^2
            ge = [E[i] <= 0 or Fraction(m) >= E2[i] for i in range(nb + 1)]
            eq_last = E[-1] >= 0 and Fraction(m) == E[-1] * E[-1]
        idx = sum(ge)
        if eq_last:
            idx = nb
        if 1 <= idx <= nb:
            counts[idx - 1] += c
    return counts


@pytest.mark.parametrize("edges,rounded", [
    (np.arange(0.0, 9.0), False),        # integer edges: ties at perfect squares, k = 0 included
    (np.arange(0.5, 9.5), False),        # half-integer: no ties
    (np.array([1.0, 1.4142135623730951, 2.0, 3.0, 3.0, 4.5, 7.0]), True),   # fl(sqrt 2) edge, duplicate
    (np.array([-1.0, 0.0, 0.25, 0.5, 13.0]), False),              # negative, narrow, beyond kmax
])
def test_mode_counts_exact_lattice(edges, rounded):
    n, box = 12, 2 * np.pi        # kF = 1: grid-unit edges are exactly the given numbers
    d = np.random.default_rng(0).standard_normal((n, n, n))
    k3, P, Nm, ex = R.powspec_vec(d, box, edges * (2 * np.pi / box), full=True)
    np.testing.assert_array_equal(ex["counts"], _lattice_counts_exact(n, edges, rounded))


def test_rounded_sqrt_edge_differs_from_exact():
    # the float histogram puts m = 2 into the bin starting at fl(sqrt 2) > sqrt 2
    edges = np.array([1.0, 1.4142135623730951, 2.0])
    assert not np.array_equal(_lattice_counts_exact(8, edges, False), _lattice_counts_exact(8, edges, True))


def test_first_fundamental_bin_has_17_modes():
    # m = 1 (5 stored), m = 2 (8), m = 3 (4): r3 counts restricted to kz >= 0
    for n in (8, 16, 32):
        d = np.random.default_rng(n).standard_normal((n, n, n))
        k, P, Nm = R.powspec_vec_fundamental(d, 100.0)
        assert Nm[0] == 17
        assert len(Nm) == math.isqrt(3 * (n // 2) ** 2)
        assert Nm.sum() == n * n * (n // 2 + 1) - 1          # every mode but DC


def test_fundamental_index_is_isqrt():
    n = 32
    d = np.random.default_rng(2).standard_normal((n, n, n))
    _, _, _, ex = R.powspec_vec_fundamental(d, 10.0, full=True)
    m = R.spectrum_power(d, p=0)[3]
    want = np.array([math.isqrt(int(v)) - 1 for v in m])
    np.testing.assert_array_equal(ex["slot"], want)


@pytest.mark.parametrize("q", [(2, 1, 3), (0, 0, 4), (3, -2, 0), (5, 0, 1)])
def test_plane_wave_multipoles(q):
    n, box, A = 16, 250.0, 0.7
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    d = A * np.cos(2 * np.pi * (q[0] * X + q[1] * Y + q[2] * Z) / n)
    kq = math.sqrt(sum(v * v for v in q))
    mu = q[2] / kq
    kF = 2 * np.pi / box
    edges = kF * np.arange(0.5, 9.5)
    k3, P, Nm, ex = R.powspec_vec(d, box, edges, p=2, full=True)
    b = int(kq - 0.5)
    W = np.prod([(np.pi * v / n / np.sin(np.pi * v / n)) ** 4 if v else 1.0 for v in q])
    stored = 2 if q[2] == 0 else 1                       # kz = 0 plane stores both +q and -q
    S0 = stored * (A * n ** 3 / 2) ** 2 * W
    norm = (box / n ** 2) ** 3
    P0 = S0 / Nm[b] * norm
    L2 = (3 * mu ** 2 - 1) / 2
    L4 = (35 * mu ** 4 - 30 * mu ** 2 + 3) / 8
    np.testing.assert_allclose(P[b], [P0, 5 * P0 * L2, 9 * P0 * L4], rtol=1e-9, atol=1e-9 * P0)
    others = np.delete(P[:, 0], b)
    assert np.nanmax(np.abs(others)) < 1e-12 * P0


def test_empty_bins_are_nan():
    n, box = 8, 2 * np.pi
    d = np.random.default_rng(4).standard_normal((n, n, n))
    k3, P, Nm = R.powspec_vec(d, box, np.array([1.1, 1.2, 1.3, 2.0]))
    assert Nm[0] == 0 and np.all(np.isnan(P[0]))
    assert np.all(np.isfinite(P[2]))


def test_bin_centres():
    box = 300.0
    e = np.array([0.01, 0.02, 0.05])
    k3, P, Nm = R.powspec_vec(np.ones((8, 8, 8)), box, e)
    np.testing.assert_allclose(k3, [0.015, 0.035])


@pytest.mark.parametrize("stencil,p", [("CIC", 2), ("TSC", 3)])
def test_shot_noise_of_uniform_catalogue(stencil, p):
    # E[P(k)] = (1/nbar) sum_n W^2(k + 2 k_N n) / W^2(k): CIC prod(1 - 2/3 s^2),
    # TSC prod(1 - s^2 + 2/15 s^4), s = sin(pi k_a / n)
    n, box, npart = 32, 1000.0, 400_000
    rng = np.random.default_rng(123)
    pos = (rng.random((npart, 3)) * box).astype(F32)
    k, P, Nm, ex, mesh = R.estimate(*pos.T, None, box, n, 2 * np.pi / box * np.arange(1, 16), stencil=stencil,
                                    p=p)
    ki = R.wavenumbers(n)
    kx, ky, kz = R.half_spectrum_grids(n)

    def alias(kk):
        s2 = np.sin(np.pi * kk / n) ** 2
        return 1 - 2 / 3 * s2 if stencil == "CIC" else 1 - s2 + 2 / 15 * s2 * s2

    def w2(kk):   # |W(k)|^2 per axis: sinc^(2p), sinc(t) = sin(pi t) / (pi t)
        return np.sinc(kk / n) ** (2 * p)

    expect = (box ** 3 / npart) * alias(kx) * alias(ky) * alias(kz) / (w2(kx) * w2(ky) * w2(kz))
    m = (kx * kx + ky * ky + kz * kz)
    slot = R.jnp_histogram_slots(np.sqrt(m.astype(float)).reshape(-1), np.arange(1, 16, dtype=float))
    e_bin = np.bincount(slot[slot >= 0], weights=np.broadcast_to(expect, m.shape).reshape(-1)[slot >= 0],
                        minlength=14)[:14] / Nm
    ratio = P[:, 0] / e_bin
    # mode-weighted mean over ~60k modes (sigma ~ 0.6 %); per bin 5 sigma with
    # sigma ~ 1 / sqrt(N/2) (the kz = 0 plane stores conjugate pairs)
    assert abs(np.sum(ratio * Nm) / np.sum(Nm) - 1) < 0.03
    big = Nm >= 200
    assert np.all(np.abs(ratio[big] - 1) < 5 / np.sqrt(Nm[big] / 2))
    # the wrong exponent p -> p-1 would leave ratio ~ 0.8 in the last bin
    assert abs(ratio[-1] - 1) < 0.1


def test_legendre_weights_exact():
    rng = np.random.default_rng(8)
    kz = rng.integers(0, 513, 2000)
    kxy = rng.integers(0, 2 * 512 ** 2, 2000)
    m = kz * kz + kxy
    L2, L4 = R.legendre_weights(kz * kz, m)
    for a, b, l2, l4 in zip(kz[:300], m[:300], L2[:300], L4[:300]):
        if b == 0:
            assert (l2, l4) == (-0.5, 0.375)
            continue
        mu2 = Fraction(int(a) * int(a), int(b))
        assert l2 == pytest.approx(float((3 * mu2 - 1) / 2), rel=1e-15, abs=0)
        assert l4 == pytest.approx(float((35 * mu2 * mu2 - 30 * mu2 + 3) / 8), rel=1e-15, abs=1e-300)
    # 3 kz^2 = m (mu^2 = 1/3, e.g. the (1, 1, 1) and (16, 16, 16) modes): L2 is exactly 0
    L2, _ = R.legendre_weights(np.array([1, 256, 3 * 49]), np.array([3, 768, 3 * 147]))
    np.testing.assert_array_equal(L2, [0.0, 0.0, 0.0])
