"""GPU parity of the HIP path (libpowspec_amd.so, through the C ABI) against the CPU oracle.

Bars: mode counts and bin indices bit-exact (int64 equality); P_l within 1e-5 of the
same-bin absolute scale A_l,b = (2l+1) sum_b |L_l(mu)| P2 norm / N_b (P_2, P_4 cross zero),
in every bin (beyond k_N too);
meshes within float32 accumulation error; FFT within float32 rounding.
"""
import math

import numpy as np
import pytest
import torch

from jax_powspec_amd import binning, correlations, mas
from oracle import ref_fp64 as R

pytestmark = pytest.mark.gpu
F32 = np.float32
PK_TOL = 1e-5


def catalogue(rng, npart, box, clumps=True, edges=True):
    pts = rng.random((npart, 3)) * box
    if clumps:   # some clustering so P_2, P_4 are not pure noise
        c = rng.random((20, 3)) * box
        k = npart // 4
        pts[:k] = (c[rng.integers(0, 20, k)] + rng.normal(0, box / 40, (k, 3))) % box
        pts[: k // 2, 2] = (pts[: k // 2, 2] + rng.normal(0, box / 20, k // 2)) % box   # LOS stretch
    if edges:
        pts[:6] = [[0, 0, 0], [box * (1 - 1e-7), 1, 2], [3, box * (1 - 1e-7), 4], [5, 6, box - 1e-3],
                   [box / 2, box / 2, box / 2], [0, box / 3, 0]]
    return pts.astype(F32)


def ref_catalogue(rng, npart, box, n):
    """Adds particles outside [0, box) that exercise the reference's drop / wrap rules."""
    p = catalogue(rng, npart, box)
    H = box / n
    extra = np.array([[box, 1, 1], [-0.4 * H, 2, 3], [-1.6 * H, 4, 5], [1.5 * box, 6, 7], [2 * box + 3, 8, 9],
                      [1, -2.5 * H, box], [box * 1.0, box * 1.0, box * 1.0], [-box - 2 * H, 1, 1]], F32)
    return np.concatenate([p, extra]).astype(F32)


def mesh_close(g, o):
    g = g.detach().cpu().numpy().astype(np.float64)
    scale = np.max(np.abs(o))
    err = np.abs(g - o)
    assert np.all(err <= 4e-6 * scale + 2e-6 * np.abs(o)), (err.max(), scale)


def pk_check(pk, nm, ex, o_pk, o_extra, box, n, shot=0.0, sum_abs=None):
    """Gate on the GPU's FP64 sums; then check the float32 outputs are their rounding.

    Bins holding only the k = 0 mode carry |sum delta|^2, a cancellation whose FP64 value
    depends on summation order; they are compared as |sqrt S_gpu - sqrt S_o| <=
    1e-12 sum|delta| instead of relatively.
    """
    cnt = ex["counts"].cpu().numpy()
    np.testing.assert_array_equal(cnt, o_extra["counts"])
    np.testing.assert_array_equal(nm.cpu().numpy(), o_extra["counts"].astype(np.float32))
    S = ex["sums"].cpu().numpy()
    norm = (box / n ** 2) ** 3
    with np.errstate(invalid="ignore", divide="ignore"):
        P = S / cnt[:, None] * norm * np.array([1.0, 5.0, 9.0])
    o = np.array(o_pk, np.float64)
    empty = cnt == 0
    assert np.all(np.isnan(pk.cpu().numpy()[empty]))
    check = ~empty
    slot0 = int(o_extra["slot"][0])            # bin of the k = 0 mode (-1: dropped)
    if slot0 >= 0 and cnt[slot0] == 1:
        check[slot0] = False
        So = o[slot0, 0] / norm
        assert abs(np.sqrt(max(S[slot0, 0], 0)) - np.sqrt(max(So, 0))) <= 1e-12 * (sum_abs or 1.0)
    scale = o_extra["scale"][check]
    err = np.abs(P[check] - o[check]) / np.maximum(scale, 1e-300)
    # every bin at 1e-5, also beyond the Nyquist sphere (k > k_N = n/2 k_F), where the
    # deconvolution exposes the float32 mesh and FFT round-off most (reported separately)
    kbin = o_extra["kmean"][check] / (2 * np.pi / box)
    corner = np.isfinite(kbin) & (kbin > n / 2)
    worst = float(np.nanmax(err / PK_TOL)) if err.size else 0.0
    below = float(np.nanmax(err[~corner])) if (~corner).any() else 0.0
    above = float(np.nanmax(err[corner])) if corner.any() else 0.0
    print(f"[pk_check n={n}] max eps (same-bin scale): k <= k_N {below:.2e}, k > k_N {above:.2e}", flush=True)
    assert worst <= 1.0, (worst, np.unravel_index(np.nanargmax(err), err.shape))
    # float32 outputs = rounding of (S / N * norm * (2l+1), minus the shot noise on l = 0)
    want32 = (P - np.array([shot, 0.0, 0.0])).astype(F32)
    np.testing.assert_allclose(pk.cpu().numpy()[~empty], want32[~empty], rtol=2e-7, atol=0)
    return worst


# ----------------------------------------------------------------------------------------
# mass assignment
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [24, 32])
@pytest.mark.parametrize("tile", [None, (1, 4, 0), (1, 2, 0), "direct"])
@pytest.mark.parametrize("fn,wrap", [("vec", True), ("vec", False), ("scan", True), ("scan", False)])
def test_reference_stencil_parity(gpu, n, tile, fn, wrap):
    rng = np.random.default_rng(100 + n)
    box = 200.0
    p = ref_catalogue(rng, 5000, box, n)
    w = (0.5 + rng.random(len(p))).astype(F32)
    base = rng.standard_normal((n, n, n)).astype(F32)
    t = torch.as_tensor(p, device=gpu)
    tw = torch.as_tensor(w, device=gpu)
    tb = torch.as_tensor(base, device=gpu)
    mode = {("vec", True): 0, ("vec", False): 1, ("scan", True): 2, ("scan", False): 3}[(fn, wrap)]
    got = mas._scatter(tb, t[:, 0], t[:, 1], t[:, 2], tw, 0.0, 0.0, 0.0, box, n, "CIC_REF", mode, tile=tile)
    want = R._accumulate(base.astype(np.float64), n,
                         R.cic_ref_deposits(p[:, 0], p[:, 1], p[:, 2], w, 0.0, 0.0, 0.0, box, n, mode))
    mesh_close(got, want)
    np.testing.assert_array_equal(tb.cpu().numpy(), base)      # functional: input untouched


def test_dropin_signatures(gpu):
    rng = np.random.default_rng(5)
    n, box = 16, 64.0
    p = catalogue(rng, 3000, box)
    t = torch.as_tensor(p, device=gpu)
    w = torch.ones(len(p), device=gpu)
    z = torch.zeros((n, n, n), device=gpu)
    for f, of in ((mas.cic_mas_vec, R.cic_mas_vec), (mas.cic_mas, R.cic_mas)):
        for wrap in (True, False):
            got = f(z, t[:, 0], t[:, 1], t[:, 2], w, len(p), 0.0, 0.0, 0.0, box, n, wrap)
            want = of(np.zeros((n,) * 3), p[:, 0], p[:, 1], p[:, 2], np.ones(len(p), F32), len(p), 0., 0., 0., box,
                      n, wrap)
            mesh_close(got, want)
    # xmin = particles[:, i].min() as tests/correlations.py:45 does
    mins = [float(p[:, i].min()) for i in range(3)]
    got = mas.cic_mas_vec(z, t[:, 0], t[:, 1], t[:, 2], w, len(p), *mins, box, n, True)
    want = R.cic_mas_vec(np.zeros((n,) * 3), *p.T, np.ones(len(p), F32), len(p), *mins, box, n, True)
    mesh_close(got, want)


@pytest.mark.parametrize("stencil", ["CIC", "TSC"])
@pytest.mark.parametrize("n,tile", [(16, None), (24, (1, 4, 0)), (32, (1, 8, 32)), (48, "direct"), (32, (1, 2, 0)),
                                    (12, (1, 6, 0))])
def test_cic_tsc_parity(gpu, stencil, n, tile):
    rng = np.random.default_rng(7 + n)
    box = 300.0
    p = catalogue(rng, 20000, box)
    w = (0.2 + rng.random(len(p))).astype(F32)
    t = torch.as_tensor(p, device=gpu)
    stats = torch.zeros(4, dtype=torch.float64, device=gpu)
    got = mas.assign(t[:, 0], t[:, 1], t[:, 2], torch.as_tensor(w, device=gpu), box_size=box, n_bins=n, mas=stencil,
                     tile=tile, stats=stats)
    want = (R.cic_mesh if stencil == "CIC" else R.tsc_mesh)(*p.T, w, 0, 0, 0, box, n)
    mesh_close(got, want)
    m = stats.cpu().numpy()
    assert m[0] == pytest.approx(float(w.astype(np.float64).sum()), rel=1e-6)
    assert m[1] == pytest.approx(n ** 3 / m[0], rel=1e-12)


def test_strided_aos_positions_and_dense_sorted_catalogue(gpu):
    # (N, 3) array columns read in place (stride 3); a dense, cell-ordered catalogue
    n, box = 16, 16.0
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    p = (np.repeat(g, 12, axis=0) + np.random.default_rng(0).random((n ** 3 * 12, 3))).astype(F32)
    t = torch.as_tensor(p, device=gpu)
    assert t[:, 1].stride(0) == 3
    for st in ("CIC", "TSC", "CIC_REF"):
        got = mas.assign(t[:, 0], t[:, 1], t[:, 2], box_size=box, n_bins=n, mas=st)
        want = R.assign(st, *p.T, None, 0, 0, 0, box, n)
        mesh_close(got, want)


# ----------------------------------------------------------------------------------------
# FFT
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256])
def test_rfftn_parity(gpu, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n)).astype(F32) + F32(0.3)
    got = correlations.rfftn(torch.as_tensor(a, device=gpu)).cpu().numpy().astype(np.complex128)
    want = np.fft.rfftn(a.astype(np.float64))
    rel2 = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel2 <= 1e-7 * max(1.0, math.log2(n ** 3)), rel2
    assert np.max(np.abs(got - want)) <= 2e-6 * np.max(np.abs(want))


def test_rfftn_plane_wave_exact_bins(gpu):
    n = 64
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    a = np.cos(2 * np.pi * (3 * X - 5 * Y + 7 * Z) / n).astype(F32)
    got = correlations.rfftn(torch.as_tensor(a, device=gpu)).cpu().numpy()
    peak = n ** 3 / 2
    assert abs(got[3, n - 5, 7] - peak) <= 1e-5 * peak
    got[3, n - 5, 7] = 0
    assert np.max(np.abs(got)) <= 2e-5 * peak


# ----------------------------------------------------------------------------------------
# P_l from a mesh
# ----------------------------------------------------------------------------------------
def _delta(rng, n, box):
    p = catalogue(rng, 40 * n ** 2, box)
    return R.contrast(R.cic_mesh(*p.T, np.ones(len(p), F32), 0, 0, 0, box, n)).astype(F32)


EDGES = {
    "kf_half": lambda n, kF: kF * (0.5 + np.arange(0, n // 2 + 1)),
    "kf_integer_with_zero": lambda n, kF: kF * np.arange(0, n // 2 + 1),
    "arbitrary": lambda n, kF: np.linspace(0.3 * kF, 0.9 * n / 2 * kF, 11),
    "narrow_global_hist": lambda n, kF: np.arange(0.2 * kF, n / 2 * kF, 0.05 * kF),   # > 256 bins
    "beyond_nyquist": lambda n, kF: np.arange(1e-4, 3 * n * kF, 1.7 * kF),
}


@pytest.mark.parametrize("n", [16, 24, 32, 64])
@pytest.mark.parametrize("edges", sorted(EDGES))
@pytest.mark.parametrize("window", ["CIC", "TSC"])
def test_powspec_vec_parity(gpu, n, edges, window):
    rng = np.random.default_rng(n * 7)
    box = 750.0
    d = _delta(rng, n, box)
    kF = 2 * np.pi / box
    k_edges = EDGES[edges](n, kF)
    k3, pk, nm, ex = correlations.powspec_vec(torch.as_tensor(d, device=gpu), box, k_edges, mas=window, full=True)
    p = 2 if window == "CIC" else 3
    ok3, opk, onm, oex = R.powspec_vec(d, box, k_edges, p=p, full=True)
    pk_check(pk, nm, ex, opk, oex, box, n, sum_abs=float(np.abs(d.astype(np.float64)).sum()))
    np.testing.assert_allclose(k3.cpu().numpy(), ok3, rtol=1e-6)
    km = ex["kmean"].cpu().numpy()
    fin = oex["counts"] > 0
    np.testing.assert_allclose(km[fin], oex["kmean"][fin], rtol=1e-6)


@pytest.mark.parametrize("n,edges", [(96, "kf_half"), (128, "kf_half"), (192, "kf_half"), (256, "kf_half"),
                                     (384, "kf_half"), (512, "kf_half"), (128, "narrow_global_hist"),
                                     (192, "narrow_global_hist")])
def test_powspec_vec_parity_mid_sizes(gpu, n, edges):
    # the x pass's strip layouts above n = 64: two 32-column strips (96, 128; register
    # prefetch at 128), two 8-column strips whose half-line partners are XCD-paired (192,
    # 256, 384, 512), and the per-block histogram in HBM for > 256 bins
    rng = np.random.default_rng(n + 5)
    box = 1200.0
    d = _delta(rng, n, box)
    kF = 2 * np.pi / box
    k_edges = EDGES[edges](n, kF)
    k3, pk, nm, ex = correlations.powspec_vec(torch.as_tensor(d, device=gpu), box, k_edges, mas="CIC", full=True)
    ok3, opk, onm, oex = R.powspec_vec(d, box, k_edges, p=2, full=True)
    pk_check(pk, nm, ex, opk, oex, box, n, sum_abs=float(np.abs(d.astype(np.float64)).sum()))


@pytest.mark.parametrize("n", [24, 96, 128, 256])
@pytest.mark.parametrize("kmax_frac,nbins", [(0.12, 7), (0.45, 40), (0.3, 1), (1.74, 90)])
def test_powspec_vec_parity_bins_end_early(gpu, n, kmax_frac, nbins):
    # the binning x pass loads only the tiles holding a mode below the last edge: bins that
    # end at a fraction of k_N leave most tiles and whole kz planes out (a single bin too),
    # and edges beyond the corner of the cube keep every tile.  Sizes: 8-column tiles
    # without strips (24), 32-column strips (96, 128) and 8-column strip pairs (256).
    rng = np.random.default_rng(n + 17 * nbins)
    box = 1000.0
    d = _delta(rng, n, box)
    kF = 2 * np.pi / box
    k_edges = np.linspace(0.5 * kF, kmax_frac * (n / 2) * kF, nbins + 1)
    print(f"[bins end early] n={n} kmax={kmax_frac} k_N nb={nbins}", flush=True)
    k3, pk, nm, ex = correlations.powspec_vec(torch.as_tensor(d, device=gpu), box, k_edges, mas="CIC", full=True)
    ok3, opk, onm, oex = R.powspec_vec(d, box, k_edges, p=2, full=True)
    pk_check(pk, nm, ex, opk, oex, box, n, sum_abs=float(np.abs(d.astype(np.float64)).sum()))
    np.testing.assert_array_equal(ex["counts"].cpu().numpy(), oracle_counts(n, k_edges, box))


@pytest.mark.parametrize("n", [16, 48, 64])
def test_powspec_vec_fundamental_parity(gpu, n):
    rng = np.random.default_rng(n)
    box = 1000.0
    d = _delta(rng, n, box)
    k3, pk, nm, ex = correlations.powspec_vec_fundamental(torch.as_tensor(d, device=gpu), box, full=True)
    ok3, opk, onm, oex = R.powspec_vec_fundamental(d, box, full=True)
    pk_check(pk, nm, ex, opk, oex, box, n)
    np.testing.assert_allclose(k3.cpu().numpy(), ok3, rtol=1e-6)


# ----------------------------------------------------------------------------------------
# whole path
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("stencil", ["CIC_REF", "CIC", "TSC"])
@pytest.mark.parametrize("n,kind,tile", [(32, "kf_half", None), (48, "fundamental", None),
                                         (64, "kf_integer_with_zero", None), (32, "kf_half", "direct"),
                                         (24, "fundamental", (1, 6, 0)), (16, "kf_half", None)])
def test_estimate_parity(gpu, stencil, n, kind, tile):
    rng = np.random.default_rng(n + len(stencil))
    box = 1000.0
    p = catalogue(rng, 60000, box)
    kF = 2 * np.pi / box
    k_edges = None if kind == "fundamental" else EDGES[kind](n, kF)
    shot = box ** 3 / len(p)
    t = torch.as_tensor(p, device=gpu)
    k3, pk, nm, ex = correlations.estimate(t[:, 0], t[:, 1], t[:, 2], box_size=box, n_bins=n, k_edges=k_edges,
                                           mas=stencil, shot_noise=shot, full=True, tile=tile)
    window = 3 if stencil == "TSC" else 2
    res = R.estimate(*p.T, None, box, n, k_edges, stencil=stencil, p=window, shot_noise=shot)
    ok3, opk, onm, oex = res[:4]
    # gate before the shot-noise subtraction; the float32 output includes it
    pk_check(pk, nm, ex, opk + [shot, 0, 0], oex, box, n, shot=shot)


@pytest.mark.parametrize("stencil,n,npart,kind", [("CIC_REF", 128, 400_000, "kf_half"), ("TSC", 128, 400_000, "fundamental"),
                                                  ("CIC", 256, 1_000_000, "kf_half"),
                                                  ("CIC_REF", 256, 1_000_000, "narrow_global_hist")])
def test_estimate_parity_mid_sizes(gpu, stencil, n, npart, kind):
    # the fused deposit with several row bundles per x row and up to 1e6 particles
    rng = np.random.default_rng(n + npart % 97)
    box = 1500.0
    p = catalogue(rng, npart, box)
    kF = 2 * np.pi / box
    k_edges = None if kind == "fundamental" else EDGES[kind](n, kF)
    shot = box ** 3 / len(p)
    t = torch.as_tensor(p, device=gpu)
    k3, pk, nm, ex = correlations.estimate(t[:, 0], t[:, 1], t[:, 2], box_size=box, n_bins=n, k_edges=k_edges,
                                           mas=stencil, shot_noise=shot, full=True)
    window = 3 if stencil == "TSC" else 2
    res = R.estimate(*p.T, None, box, n, k_edges, stencil=stencil, p=window, shot_noise=shot)
    pk_check(pk, nm, ex, res[1] + [shot, 0, 0], res[3], box, n, shot=shot)


def test_estimate_equals_two_step_path(gpu):
    # estimate() == assign + (delta / mean - 1 in torch) + powspec_vec, at 128^3
    rng = np.random.default_rng(42)
    n, box = 128, 1000.0
    p = torch.as_tensor(catalogue(rng, 1_000_000, box), device=gpu)
    kF = 2 * np.pi / box
    edges = kF * np.arange(1, n // 2 + 1)
    k1, pk1, nm1, ex1 = correlations.estimate(p[:, 0], p[:, 1], p[:, 2], box_size=box, n_bins=n, k_edges=edges,
                                              full=True)
    m = mas.cic_mas_vec(torch.zeros((n, n, n), device=gpu), p[:, 0], p[:, 1], p[:, 2],
                        torch.ones(p.shape[0], device=gpu), p.shape[0], 0.0, 0.0, 0.0, box, n, True)
    d = (m.double() / m.double().mean() - 1.0).float()
    k2, pk2, nm2, ex2 = correlations.powspec_vec(d, box, edges, full=True)
    assert torch.equal(ex1["counts"], ex2["counts"])
    scale = (ex2["sums"].abs() / ex2["counts"].unsqueeze(1)).cpu().numpy() * (box / n ** 2) ** 3 * [1, 5, 9]
    err = np.abs(pk1.cpu().numpy() - pk2.cpu().numpy()) / scale[:, :1]
    assert np.nanmax(err) < 2e-5


def test_bin_sums_are_bitwise_reproducible(gpu, c2_mock):
    # the bins sum in 64-bit fixed point (a Parseval bound of the binned power fixes the
    # exponent), so repeated estimates agree bit for bit: the fused path on the 512^3
    # lognormal + RSD mock, and the mesh path (powspec_vec) on its assigned mesh, each twice
    box = C2_BOX
    kF = 2 * np.pi / box
    edges = kF * (0.5 + np.arange(C2_N // 2 + 1))
    x, y, z = c2_mock[:, 0], c2_mock[:, 1], c2_mock[:, 2]
    runs = [correlations.estimate(x, y, z, box_size=box, n_bins=C2_N, k_edges=edges, full=True) for _ in range(2)]
    for a, b in [(runs[0], runs[1])]:
        assert torch.equal(a[3]["sums"], b[3]["sums"])
        assert torch.equal(a[3]["counts"], b[3]["counts"])
        assert torch.equal(torch.nan_to_num(a[1]), torch.nan_to_num(b[1]))
    m = mas.assign(x, y, z, box_size=box, n_bins=C2_N, mas="CIC_REF")
    d = (m / m.mean() - 1.0).float()
    del m
    ps = [correlations.powspec_vec(d, box, edges, full=True) for _ in range(2)]
    assert torch.equal(ps[0][3]["sums"], ps[1][3]["sums"])
    assert torch.equal(ps[0][3]["counts"], ps[1][3]["counts"])
    # the two paths transform different float rounding of the same field: close, not equal
    s1, s2 = runs[0][3]["sums"].cpu().numpy(), ps[0][3]["sums"].cpu().numpy()
    scale = np.abs(s2).max(axis=0, keepdims=True)
    print(f"[bitwise] fused vs mesh path: max |diff| / column scale {np.max(np.abs(s1 - s2) / scale):.2e}",
          flush=True)


def test_non_finite_and_zero_fields(gpu):
    # a NaN or an infinity anywhere in the field makes every bin's power NaN (as FP64 sums of
    # it would be); an all-zero field has zero power in every bin (no fixed-point scale)
    n, box = 32, 200.0
    edges = (2 * np.pi / box) * (0.5 + np.arange(n // 2 + 1))
    rng = np.random.default_rng(5)
    base = torch.as_tensor(rng.standard_normal((n, n, n)).astype(np.float32), device=gpu)
    for bad in (float("nan"), float("inf")):
        d = base.clone()
        d[3, 4, 5] = bad
        k, pk, nm, ex = correlations.powspec_vec(d, box, edges, full=True)
        assert torch.isnan(pk).all(), bad
        assert torch.isnan(ex["sums"]).all(), bad
        assert torch.equal(ex["counts"], correlations.powspec_vec(base, box, edges, full=True)[3]["counts"])
    k, pk, nm, ex = correlations.powspec_vec(torch.zeros((n, n, n), device=gpu), box, edges, full=True)
    assert torch.equal(ex["sums"], torch.zeros_like(ex["sums"]))
    assert torch.isfinite(pk).all()


C2_N, C2_BOX, C2_NPART = 512, 1500.0, 20_000_000


def oracle_counts(n, k_edges, box):
    """Stored half-spectrum entries per bin under jnp.histogram's rules (FP64 k, FP64
    edges in grid units): an independent count over all n^2 (n/2 + 1) entries."""
    kx, ky, kz = R.half_spectrum_grids(n)
    m = (kx * kx + ky * ky + kz * kz).astype(np.int64).reshape(-1)
    e = np.asarray(k_edges, np.float64) / (2 * np.pi / box)
    slot = R.jnp_histogram_slots(np.sqrt(m.astype(np.float64)), e)
    return np.bincount(slot[slot >= 0], minlength=e.size - 1)[: e.size - 1].astype(np.int64)


@pytest.fixture(scope="module")
def c2_mock(gpu):
    """512^3 lognormal + redshift-space mock from the GPU generator (the headline spectrum
    P(k) = 2e4 (k / 0.05)^-1.5 exp(-(k / 0.5)^2), nbar = 2e7 / L^3, growth rate 0.7)."""
    from jax_powspec_amd import gauss_field
    kf = np.geomspace(1e-4, 10.0, 2000)
    pkf = 2e4 * (kf / 0.05) ** -1.5 * np.exp(-(kf / 0.5) ** 2)
    pos = gauss_field.lognormal_mock(C2_N, kf, pkf, C2_BOX, C2_NPART / C2_BOX ** 3, 1.0, 2002, growth_rate=0.7,
                                     device=gpu)
    torch.cuda.synchronize()
    return pos


@pytest.mark.timeout(300)
@pytest.mark.parametrize("edges", ["kf_half", "kf_integer"])
def test_full_size_device_counts(gpu, c2_mock, edges):
    # 512^3, ~2e7 lognormal + RSD particles: the per-bin counts the binning pass itself
    # produces equal an independent FP64 jnp.histogram count of every stored entry, for
    # half-integer edges (no ties) and integer edges (every m = k^2 on an edge)
    n, box = C2_N, C2_BOX
    kF = 2 * np.pi / box
    k_edges = kF * (0.5 + np.arange(257)) if edges == "kf_half" else kF * np.arange(1, 258)
    p = c2_mock
    print(f"[device counts] n={n} N_p={p.shape[0]} edges={edges}", flush=True)
    k3, pk, nm, ex = correlations.estimate(p[:, 0], p[:, 1], p[:, 2], box_size=box, n_bins=n, k_edges=k_edges,
                                           full=True)
    want = oracle_counts(n, k_edges, box)
    np.testing.assert_array_equal(ex["counts"].cpu().numpy(), want)
    np.testing.assert_array_equal(nm.cpu().numpy(), want.astype(np.float32))
    # the host lattice tables (used only for the mean |k|) agree with both
    tab = binning.TABLES.get(gpu, n, binning.grid_edges(k_edges, box))
    np.testing.assert_array_equal(tab.modes.cpu().numpy(), want)
    # repeated runs: identical counts and identical bin sums (the binning adds 64-bit
    # fixed-point integers, whose sums do not depend on the order of the adds)
    k3b, pkb, nmb, exb = correlations.estimate(p[:, 0], p[:, 1], p[:, 2], box_size=box, n_bins=n, k_edges=k_edges,
                                               full=True)
    assert torch.equal(ex["counts"], exb["counts"])
    assert torch.equal(ex["sums"], exb["sums"])


def f32_fft_floor(mesh, box, n, k_edges, p, A):
    """Per-bin error of a float32 FFT (scipy, complex64) of the float32 rho - mean mesh,
    binned in FP64 by the oracle's rules: the round-off any float32 mesh + FFT path has."""
    import scipy.fft
    mesh = np.asarray(mesh, np.float64)
    M = mesh.sum()
    rows = (mesh - M / n ** 3).astype(F32)
    dk32 = scipy.fft.rfftn(rows, workers=16).astype(np.complex128) * (n ** 3 / M)
    dk64 = np.fft.rfftn(mesh / (M / n ** 3) - 1.0)
    out = []
    for dk in (dk32, dk64):
        dk[0, 0, 0] = 0.0
        d2 = (np.abs(dk * R.window_correction(n, p)) ** 2).reshape(-1)
        kx, ky, kz = R.half_spectrum_grids(n)
        m = (kx * kx + ky * ky + kz * kz).astype(np.int64).reshape(-1)
        kz2 = np.broadcast_to(kz * kz, (n, n, n // 2 + 1)).astype(np.int64).reshape(-1)
        slot = R.jnp_histogram_slots(np.sqrt(m.astype(np.float64)), np.asarray(k_edges) / (2 * np.pi / box))
        S0, S2, S4 = R._bin_sums(slot, len(k_edges) - 1, d2, kz2, m, np.sqrt(m.astype(np.float64)))[:3]
        out.append(np.stack([S0, S2, S4], 1))
    norm = (box / n ** 2) ** 3
    cnt = np.maximum(oracle_counts(n, k_edges, box), 1)[:, None]
    return np.abs(out[0] - out[1]) * norm * np.array([1.0, 5.0, 9.0]) / cnt / A


@pytest.mark.timeout(600)
def test_headline_catalogue_vs_fp64_oracle(gpu):
    # the measured workload itself: 512^3, ~2e7 lognormal + RSD particles (bench.py's
    # catalogue), CIC reference stencil, P_0,2,4 in 256 k_F bins, against the FP64 oracle on
    # the same float32 positions; per-multipole error stated
    import bench
    n, box = C2_N, C2_BOX
    p = bench.make_catalogue(n, box, C2_NPART / box ** 3, gpu)
    npart = p.shape[0]
    kF = 2 * np.pi / box
    k_edges = kF * (0.5 + np.arange(257))
    shot = box ** 3 / npart
    print(f"[headline] n={n} N_p={npart}: GPU estimate", flush=True)
    k3, pk, nm, ex = correlations.estimate(p[:, 0], p[:, 1], p[:, 2], box_size=box, n_bins=n, k_edges=k_edges,
                                           mas="CIC_REF", shot_noise=shot, full=True)
    torch.cuda.synchronize()
    pos = p.cpu().numpy()
    del p
    print("[headline] FP64 oracle on the host", flush=True)
    res = R.estimate(pos[:, 0], pos[:, 1], pos[:, 2], None, box, n, k_edges, stencil="CIC_REF", p=2,
                     shot_noise=shot)
    opk, oex, mesh = res[1], res[3], res[-1]
    worst = pk_check(pk, nm, ex, opk + [shot, 0, 0], oex, box, n, shot=shot)
    cnt = ex["counts"].cpu().numpy()
    S = ex["sums"].cpu().numpy()
    P = S / cnt[:, None] * (box / n ** 2) ** 3 * np.array([1.0, 5.0, 9.0])
    o = np.asarray(opk, np.float64) + [shot, 0, 0]
    eps = np.abs(P - o) / oex["scale"]
    rel = np.abs(P - o) / np.abs(o)
    floor = f32_fft_floor(mesh, box, n, k_edges, 2, oex["scale"])
    for l in range(3):
        print(f"[headline] l={2 * l}: max eps (same-bin scale) {eps[:, l].max():.2e}, "
              f"median plain relative {np.median(rel[:, l]):.2e}, float32-FFT floor {floor[:, l].max():.2e}",
              flush=True)
    assert worst <= 1.0
    assert np.all(eps <= PK_TOL), eps.max()


def test_estimate_reference_stencil_with_outside_particles(gpu):
    # irregular particles of the reference stencil (outside [0, L)) in the fused path
    rng = np.random.default_rng(77)
    n, box = 32, 400.0
    p = ref_catalogue(rng, 30000, box, n)
    kF = 2 * np.pi / box
    edges = kF * (0.5 + np.arange(n // 2 + 1))
    t = torch.as_tensor(p, device=gpu)
    for tile in (None, "direct"):
        k3, pk, nm, ex = correlations.estimate(t[:, 0], t[:, 1], t[:, 2], box_size=box, n_bins=n, k_edges=edges,
                                               full=True, tile=tile)
        ok3, opk, onm, oex = R.estimate(*p.T, None, box, n, edges, stencil="CIC_REF")[:4]
        pk_check(pk, nm, ex, opk, oex, box, n)
