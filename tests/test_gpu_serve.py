"""PkServer (continuous batching): every result equals the plain estimate() of the same
request, alone and in a concurrent mix, under a budget that forces engine eviction.

Bars: mode counts bit-exact; P_l within 3e-7 relative of the plain path (the same kernels
and tiles; only the x-pass block count, hence the FP64 partial-sum order, differs), NaN in
the same (empty) bins.
"""
import numpy as np
import pytest
import torch

from jax_powspec_amd import correlations
from jax_powspec_amd.serve import PkServer
from test_gpu_parity import catalogue

pytestmark = pytest.mark.gpu


def requests(gpu, specs, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n, npart, mas, kind in specs:
        box = float(rng.uniform(300, 1200))
        p = torch.as_tensor(catalogue(rng, npart, box, edges=False), device=gpu)
        kF = 2 * np.pi / box
        edges = None if kind == "fundamental" else kF * (0.5 + np.arange(n // 2 + 1))
        out.append(dict(x=p[:, 0], y=p[:, 1], z=p[:, 2], box_size=box, n_bins=n, mas=mas, k_edges=edges,
                        shot_noise=box ** 3 / npart))
    return out


def plain(r):
    k, pk, nm = correlations.estimate(r["x"], r["y"], r["z"], box_size=r["box_size"], n_bins=r["n_bins"],
                                      k_edges=r["k_edges"], mas=r["mas"], shot_noise=r["shot_noise"])
    return k.cpu().numpy(), pk.cpu().numpy(), nm.cpu().numpy()


def same(a, b):
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(np.isnan(a[1]), np.isnan(b[1]))
    ok = ~np.isnan(b[1])
    np.testing.assert_allclose(a[1][ok], b[1][ok], rtol=3e-7, atol=3e-7 * np.abs(b[1][ok]).max())
    np.testing.assert_allclose(a[0], b[0], rtol=1e-6)


SPECS = [(32, 20000, "CIC", "kf"), (64, 50000, "TSC", "kf"), (32, 3000, "CIC_REF", "fundamental"),
         (64, 10000, "CIC", "fundamental"), (768, 200000, "CIC", "kf"), (32, 40000, "TSC", "kf"),
         (64, 70000, "CIC_REF", "kf"), (48, 30000, "CIC", "kf")]


def test_alone_and_in_the_mix(gpu):
    reqs = requests(gpu, SPECS, seed=1)
    want = [plain(r) for r in reqs]
    srv = PkServer(device=gpu, max_batch=4, max_particles=300_000)
    try:
        alone = [srv.submit(r).result(timeout=60) for r in reqs]          # one at a time
        futs = [srv.submit(r) for r in reqs + reqs[::-1]]                 # all at once, twice
        mixed = [f.result(timeout=60) for f in futs]
        st = srv.stats()
    finally:
        srv.close()
    for a, w in zip(alone, want):
        same(a, w)
    for m, w in zip(mixed, want + want[::-1]):
        same(m, w)
    assert st["requests"] == 3 * len(reqs) and st["errors"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_budget_forces_eviction(gpu):
    reqs = requests(gpu, [(32, 20000, "CIC", "kf"), (64, 20000, "TSC", "kf")] * 3, seed=2)
    want = [plain(r) for r in reqs]
    probe = PkServer(device=gpu, max_batch=2, max_particles=20000, engines_per_n=1)
    try:
        one = probe._estimate_bytes(64)
    finally:
        probe.close()
    srv = PkServer(device=gpu, max_batch=2, max_particles=20000, engines_per_n=1, hbm_budget=int(1.2 * one))
    try:
        got = [f.result(timeout=60) for f in [srv.submit(r) for r in reqs]]
        assert srv.reserved <= srv.hbm_budget
    finally:
        srv.close()
    for g, w in zip(got, want):
        same(g, w)


def test_rejects_oversized_requests(gpu):
    reqs = requests(gpu, [(32, 5000, "CIC", "kf")], seed=3)
    srv = PkServer(device=gpu, max_particles=1000)
    try:
        with pytest.raises(ValueError):
            srv.submit(reqs[0])
    finally:
        srv.close()


@pytest.mark.timeout(300)
def test_served_results_against_oracle_with_a_1024_request(gpu):
    # a mix submitted at once (so requests overlap on the device), including a 1024^3 one;
    # the small-n results against the FP64 oracle at 1e-5 of the same-bin scale, the 1024^3
    # one against the plain path (the oracle at 1024^3 runs in test_gpu_large_mesh.py)
    from oracle import ref_fp64 as R
    specs = [(64, 40000, "CIC", "kf"), (128, 200000, "TSC", "kf"), (1024, 500000, "CIC", "kf"),
             (32, 10000, "CIC_REF", "fundamental"), (96, 80000, "TSC", "fundamental")]
    reqs = requests(gpu, specs, seed=7)
    srv = PkServer(device=gpu, max_batch=4, max_particles=600_000)
    try:
        got = [f.result(timeout=120) for f in [srv.submit(r) for r in reqs]]
        st = srv.stats()
    finally:
        srv.close()
    assert st["errors"] == 0
    for r, g in zip(reqs, got):
        n = r["n_bins"]
        print(f"[served] n={n} {r['mas']} N_p={r['x'].numel()}", flush=True)
        if n >= 1024:
            same(g, plain(r))
            continue
        pos = torch.stack([r["x"], r["y"], r["z"]], 1).cpu().numpy()
        window = 3 if r["mas"] == "TSC" else 2
        res = R.estimate(*pos.T, None, r["box_size"], n, r["k_edges"], stencil=r["mas"], p=window,
                         shot_noise=r["shot_noise"])
        opk, oex = np.asarray(res[1], np.float64), res[3]
        np.testing.assert_array_equal(g[2], oex["counts"].astype(np.float32))
        ok = oex["counts"] > 0
        assert np.all(np.isnan(g[1][~ok]))
        err = np.abs(g[1][ok].astype(np.float64) - opk[ok]) / oex["scale"][ok]
        assert np.nanmax(err) <= 1e-5, np.nanmax(err)
