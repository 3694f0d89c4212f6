"""Large meshes on the sorted fixed-point path (src/mas.py:88-153 has no mesh-size limit).

At n = 768 and 1024 the row-bundle tiles are 16 rows (their 64-bit fixed-point mesh rows
and the fused path's 32-bit cells fit LDS), and the count / place kernels' per-key tables
(73k-131k keys) live in HBM.  Checked against the FP64 oracle: the assigned mesh
(pk_scatter) and the whole fused estimate's P_0,2,4 (1e-5 of the same-bin scale, counts
exact).  The oracle's FFT runs through scipy.fft with host threads and its binning slab by
slab, using the oracle's own functions, so that 1024^3 fits the host in FP64.
"""
import numpy as np
import pytest
import scipy.fft
import torch

from jax_powspec_amd import correlations, gauss_field, mas
from oracle import ref_fp64 as R

from test_gpu_parity import PK_TOL, catalogue, mesh_close

pytestmark = pytest.mark.gpu


def oracle_pk_slabs(mesh64, box, n, k_edges, p, slab=64):
    """FP64 oracle P_l (before shot noise) of a mesh: contrast, rfftn (scipy, FP64), k = 0
    set to 0 as the fused path's contrast does, window, jnp.histogram bins and Legendre
    sums of R._bin_sums, accumulated over slabs of kx."""
    M = mesh64.sum()
    delta = mesh64 * (n ** 3 / M)
    delta -= 1.0
    dk = scipy.fft.rfftn(delta, workers=16)
    del delta
    dk[0, 0, 0] = 0.0
    e = np.asarray(k_edges, np.float64) / (2 * np.pi / box)
    nb = e.size - 1
    acc = [np.zeros(nb) for _ in range(5)]
    cnt = np.zeros(nb, np.int64)
    kx, ky, kz = R.half_spectrum_grids(n)
    w1 = lambda k: (1.0 / np.sinc(k / n)) ** p        # the oracle's per-axis window factor
    for i0 in range(0, n, slab):
        sl = slice(i0, min(n, i0 + slab))
        sx = kx[sl]
        W = w1(sx) * w1(ky) * w1(kz)
        d2 = (np.abs(dk[sl] * W) ** 2).reshape(-1)
        m = (sx * sx + ky * ky + kz * kz).astype(np.int64).reshape(-1)
        kz2 = np.broadcast_to(kz * kz, (sx.shape[0], n, n // 2 + 1)).astype(np.int64).reshape(-1)
        kk = np.sqrt(m.astype(np.float64))
        slot = R.jnp_histogram_slots(kk, e)
        S0, S2, S4, c, _, A2, A4 = R._bin_sums(slot, nb, d2, kz2, m, kk)
        for a, v in zip(acc, (S0, S2, S4, A2, A4)):
            a += v
        cnt += c
    norm = (box / n ** 2) ** 3
    with np.errstate(invalid="ignore", divide="ignore"):
        P = np.stack(acc[:3], 1) / cnt[:, None] * norm * np.array([1.0, 5.0, 9.0])
        A = np.stack([acc[0], acc[3], acc[4]], 1) / cnt[:, None] * norm * np.array([1.0, 5.0, 9.0])
    return P, A, cnt


@pytest.mark.timeout(900)
@pytest.mark.parametrize("n,stencil,npart", [(768, "TSC", 3_000_000), (1024, "CIC", 6_000_000)])
def test_large_mesh_against_fp64_oracle(gpu, n, stencil, npart):
    box = 2000.0
    kf = np.geomspace(1e-4, 10.0, 2000)
    pkf = 2e4 * (kf / 0.05) ** -1.5 * np.exp(-(kf / 0.5) ** 2)
    print(f"[large] n={n} {stencil}: lognormal mock", flush=True)
    p = gauss_field.lognormal_mock(n, kf, pkf, box, npart / box ** 3, 1.0, 7000 + n, growth_rate=0.7, device=gpu)
    kF = 2 * np.pi / box
    k_edges = kF * (0.5 + 2.0 * np.arange(n // 4 + 1))      # 2 k_F wide bins up to k_N
    window = 3 if stencil == "TSC" else 2
    # the whole fused estimate (sorted, 32-bit cells, HBM count / place tables)
    k3, pk, nm, ex = correlations.estimate(p[:, 0], p[:, 1], p[:, 2], box_size=box, n_bins=n, k_edges=k_edges,
                                           mas=stencil, full=True)
    counts = ex["counts"].cpu().numpy()
    S = ex["sums"].cpu().numpy()
    # the assigned mesh of pk_scatter (sorted, 64-bit cells), then freed on the device
    got_mesh = mas.assign(p[:, 0], p[:, 1], p[:, 2], box_size=box, n_bins=n, mas=stencil)
    gm = got_mesh.cpu().numpy()
    del got_mesh
    pos = p.cpu().numpy()
    del p
    torch.cuda.empty_cache()
    print(f"[large] n={n}: FP64 oracle mesh ({pos.shape[0]} particles)", flush=True)
    mesh64 = R.assign(stencil, pos[:, 0], pos[:, 1], pos[:, 2], None, 0.0, 0.0, 0.0, box, n)
    err = np.abs(gm.astype(np.float64) - mesh64)
    scale = np.max(np.abs(mesh64))
    assert np.all(err <= 4e-6 * scale + 2e-6 * np.abs(mesh64)), (err.max(), scale)
    del gm, err
    print(f"[large] n={n}: FP64 oracle P_l", flush=True)
    P_o, A_o, cnt_o = oracle_pk_slabs(mesh64, box, n, k_edges, window)
    np.testing.assert_array_equal(counts, cnt_o)
    P = S / counts[:, None] * (box / n ** 2) ** 3 * np.array([1.0, 5.0, 9.0])
    eps = np.abs(P - P_o) / A_o
    print(f"[large] n={n}: max eps (same-bin scale) P0 {eps[:, 0].max():.2e} P2 {eps[:, 1].max():.2e} "
          f"P4 {eps[:, 2].max():.2e}", flush=True)
    assert np.all(eps <= PK_TOL), eps.max()


@pytest.mark.parametrize("stencil", ["CIC_REF", "CIC", "TSC"])
def test_hbm_tables_with_small_tiles(gpu, stencil):
    # 2-row tiles at 256^3: 65537+ keys, so count and place keep their tables in HBM (global
    # atomics on each block's row of the count matrix); the mesh against the oracle
    n, box = 256, 500.0
    rng = np.random.default_rng(11)
    pts = catalogue(rng, 400_000, box)
    t = torch.as_tensor(pts, device=gpu)
    got = mas.assign(t[:, 0], t[:, 1], t[:, 2], box_size=box, n_bins=n, mas=stencil, tile=(1, 2, 0))
    want = R.assign(stencil, *pts.T, None, 0, 0, 0, box, n)
    mesh_close(got, want)
