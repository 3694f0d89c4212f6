"""Batched estimates (pk_batch_* through jax_powspec_amd.batch) against the plain path and the oracle.

Bars: mode counts bit-exact; per-bin sums bit-identical to the plain pk_estimate's (same
kernels and tile; the x-pass block count differs, but the bins sum in 64-bit fixed point,
so the order of the partial sums does not matter); float32 outputs within 2 ulp; oracle
parity as the plain path (1e-5 of the same-bin scale).  Graph replays are checked one by one against fresh plain
results, with eager torch work and synchronisations between them.
"""
import numpy as np
import pytest
import torch

from jax_powspec_amd import batch, correlations
from oracle import ref_fp64 as R
from test_gpu_parity import catalogue, pk_check

pytestmark = pytest.mark.gpu
F32 = np.float32


def make_cats(gpu, specs, seed=0):
    """specs: (n, npart, mas, weighted, edges_kind) -> catalogue dicts (+ host copies)."""
    rng = np.random.default_rng(seed)
    cats, host = [], []
    for i, (n, npart, mas, weighted, kind) in enumerate(specs):
        box = float(rng.uniform(500, 1500))
        p = catalogue(rng, npart, box, edges=npart >= 10)
        w = (0.5 + rng.random(npart)).astype(F32) if weighted else None
        kF = 2 * np.pi / box
        edges = None if kind == "fundamental" else kF * np.arange(1, n // 2 + 1)
        t = torch.as_tensor(p, device=gpu)
        c = dict(x=t[:, 0], y=t[:, 1], z=t[:, 2], w=None if w is None else torch.as_tensor(w, device=gpu),
                 box_size=box, n_bins=n, mas=mas, k_edges=edges, shot_noise=box ** 3 / max(npart, 1))
        cats.append(c)
        host.append((p, w))
    return cats, host


def plain(c, tile):
    return correlations.estimate(c["x"], c["y"], c["z"], c.get("w"), box_size=c["box_size"], n_bins=c["n_bins"],
                                 k_edges=c["k_edges"], mas=c["mas"], shot_noise=c["shot_noise"], full=True, tile=tile)


def same_result(got, want):
    k1, pk1, nm1, ex1 = got
    k2, pk2, nm2, ex2 = want
    assert torch.equal(ex1["counts"], ex2["counts"])
    assert torch.equal(nm1, nm2)
    assert torch.equal(ex1["sums"], ex2["sums"])
    a, b = pk1.cpu().numpy(), pk2.cpu().numpy()
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    np.testing.assert_allclose(a[ok], b[ok], rtol=3e-7, atol=1e-30 + 3e-7 * np.abs(b[ok]).max())
    np.testing.assert_allclose(k1.cpu().numpy(), k2.cpu().numpy(), rtol=1e-6)


SPECS = [
    (32, 20000, "CIC", False, "kf"),
    (32, 5000, "TSC", True, "kf"),
    (32, 40000, "CIC_REF", False, "fundamental"),
    (32, 7, "TSC", False, "kf"),          # a near-empty catalogue
    (64, 60000, "CIC", False, "kf"),
    (64, 3000, "TSC", False, "fundamental"),
    (64, 100000, "CIC_REF", True, "kf"),
    (48, 30000, "TSC", False, "kf"),
]


def test_batched_equals_plain(gpu):
    cats, _ = make_cats(gpu, SPECS, seed=1)
    got = batch.powspec_batched(cats, tile=None, full=True)
    for c, g in zip(cats, got):
        same_result(g, plain(c, None))


@pytest.mark.parametrize("n", [64, 96])
def test_batched_alt_tile_equals_plain(gpu, n):
    cats, _ = make_cats(gpu, [(n, 50000, "CIC", False, "kf"), (n, 20000, "TSC", False, "kf"),
                              (n, 80000, "CIC_REF", False, "kf")], seed=n)
    tile = (1, 16, 0)
    got = batch.powspec_batched(cats, tile=tile, full=True)
    for c, g in zip(cats, got):
        same_result(g, plain(c, tile))


def test_batched_against_oracle(gpu):
    specs = [(32, 30000, "CIC", False, "kf"), (32, 30000, "TSC", False, "kf"),
             (32, 30000, "CIC_REF", False, "kf")]
    cats, host = make_cats(gpu, specs, seed=3)
    got = batch.powspec_batched(cats, tile=None, full=True)
    for c, (p, w), g in zip(cats, host, got):
        k3, pk, nm, ex = g
        window = 3 if c["mas"] == "TSC" else 2
        res = R.estimate(*p.T, None, c["box_size"], c["n_bins"], c["k_edges"], stencil=c["mas"], p=window,
                         shot_noise=c["shot_noise"])
        ok3, opk, onm, oex = res[:4]
        pk_check(pk, nm, ex, opk + [c["shot_noise"], 0, 0], oex, c["box_size"], c["n_bins"], shot=c["shot_noise"])


def test_graph_replays_follow_new_data(gpu):
    specs = [(64, 40000, "CIC", False, "kf"), (64, 25000, "TSC", True, "kf"), (64, 60000, "CIC_REF", False, "kf")]
    cats, _ = make_cats(gpu, specs, seed=4)
    b = batch.BatchEstimator(64, len(cats), 60000, 64, graph=True, device=gpu)
    b.set(cats)
    for rep in range(4):   # rep 0: plain launches, rep 1: capture + replay, reps 2-3: replays
        if rep:
            for c in cats:   # eager work between replays: move every particle (in place)
                L = c["box_size"]
                c["z"].add_(0.37 * L * rep).remainder_(L)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        res = b.results(full=True)
        for c, g in zip(cats, res):
            same_result(g, plain(c, None))
        tmp = torch.randn(1000, device=gpu).sum()   # unrelated eager work
        torch.cuda.synchronize()
        assert torch.isfinite(tmp)


def test_refill_with_fewer_smaller_jobs(gpu):
    cats, _ = make_cats(gpu, [(32, 30000, "CIC", False, "kf")] * 3, seed=5)
    b = batch.BatchEstimator(32, 4, 30000, 16, graph=True, device=gpu)
    b.set(cats).run().run()
    small, _ = make_cats(gpu, [(32, 500, "TSC", False, "kf")], seed=6)
    b.set(small).run().run()
    torch.cuda.synchronize()
    same_result(b.result(0, full=True), plain(small[0], None))
    with pytest.raises(ValueError):
        big, _ = make_cats(gpu, [(32, 40000, "CIC", False, "kf")], seed=7)
        b.set(big)


def test_autotune_searches_tiles_and_knobs(gpu):
    cats, _ = make_cats(gpu, [(64, 30000, "CIC", False, "kf"), (64, 30000, "TSC", False, "kf")], seed=10)
    best, timings = batch.autotune(64, cats, reps=1)
    tys = {t[1] for t in batch.tile_candidates(64)}
    assert len(timings) == len(tys) * len(batch.COUNT_ROUNDS) * len(batch.PASS_PCTS) ** 2
    assert best["tile"][1] in tys and best["ms"] == min(timings.values())
    # every tuned configuration gives the plain path's results
    got = batch.powspec_batched(cats, tile="auto", full=True)
    for c, g in zip(cats, got):
        same_result(g, plain(c, best["tile"]))
    b = batch.BatchEstimator(64, 2, 30000, 32, tile=best["tile"], device=gpu)
    for knobs in [(1, 50, 50), (8, 100, 50), (2, 50, 100)]:
        b.tune(*knobs).set(cats).run()
        torch.cuda.synchronize()
        for i, c in enumerate(cats):
            same_result(b.result(i, full=True), plain(c, best["tile"]))


def test_autotune_reports_a_compiled_tile(gpu):
    cats, _ = make_cats(gpu, [(64, 30000, "CIC", False, "kf"), (64, 30000, "TSC", False, "kf")], seed=8)
    best, timings = batch.autotune_tile(64, cats, reps=2)
    assert best in batch.tile_candidates(64)
    assert set(timings) == {t[1] for t in batch.tile_candidates(64)}
    assert all(v > 0 for v in timings.values())


def test_batched_large_mesh_with_hbm_tables(gpu):
    # n = 768: 16-row tiles, and every job's count / place tables in HBM (74k-111k keys)
    cats, _ = make_cats(gpu, [(768, 200000, "CIC", False, "kf"), (768, 100000, "TSC", False, "kf")], seed=9)
    got = batch.powspec_batched(cats, tile=None, full=True)
    for c, g in zip(cats, got):
        same_result(g, plain(c, None))


def lognormal_cats(gpu, specs):
    """Lognormal + RSD mocks (the headline spectrum) from the GPU generator:
    specs (n, npart, mas, box, seed) -> catalogue dicts and host positions."""
    from jax_powspec_amd import gauss_field
    kf = np.geomspace(1e-4, 10.0, 2000)
    pkf = 2e4 * (kf / 0.05) ** -1.5 * np.exp(-(kf / 0.5) ** 2)
    cats, host = [], []
    for n, npart, mas, box, seed in specs:
        p = gauss_field.lognormal_mock(n, kf, pkf, box, npart / box ** 3, 1.0, seed, growth_rate=0.7, device=gpu)
        kF = 2 * np.pi / box
        cats.append(dict(x=p[:, 0], y=p[:, 1], z=p[:, 2], box_size=box, n_bins=n, mas=mas,
                         k_edges=kF * np.arange(1, n // 2 + 1), shot_noise=box ** 3 / p.shape[0]))
        host.append(p.cpu().numpy())
    return cats, host


@pytest.mark.timeout(600)
def test_batched_covariance_mix_against_oracle(gpu):
    # jobs shaped like the covariance set: n in {128, 192, 256}, 1e5 - 4e6 lognormal + RSD
    # particles, CIC and TSC, k_F bins (integer edges: ties), two batches by mesh side run
    # concurrently; every job against the FP64 oracle
    specs = [(128, 1e5, "CIC", 600.0, 3101), (128, 4e6, "TSC", 1500.0, 3102), (192, 1e6, "CIC", 1000.0, 3103),
             (192, 3e5, "TSC", 800.0, 3104), (256, 4e6, "CIC", 1400.0, 3105), (256, 2e6, "TSC", 1200.0, 3106)]
    cats, host = lognormal_cats(gpu, specs)
    got = batch.powspec_batched(cats, tile=None, full=True)
    for c, p, g in zip(cats, host, got):
        k3, pk, nm, ex = g
        window = 3 if c["mas"] == "TSC" else 2
        print(f"[covariance mix] n={c['n_bins']} {c['mas']} N_p={p.shape[0]}", flush=True)
        res = R.estimate(*p.T, None, c["box_size"], c["n_bins"], c["k_edges"], stencil=c["mas"], p=window,
                         shot_noise=c["shot_noise"])
        pk_check(pk, nm, ex, res[1] + [c["shot_noise"], 0, 0], res[3], c["box_size"], c["n_bins"],
                 shot=c["shot_noise"])
