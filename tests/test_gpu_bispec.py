"""GPU parity of the bispectrum (pk_bispec_mesh through the C ABI) against the FP64 oracle
(oracle/ref_fp64.py: bispec).

Bars: P_i within 1e-5 relative (sums of squares); B within 2e-5 of its absolute scale
sum |d1 d2 d3| / sum I1 I2 I3 (L^2/n^3)^3 (float32 FFTs of the three shell fields); Q within
the propagated bound; k_all and theta equal to the host computation.
"""
import numpy as np
import pytest
import torch

from jax_powspec_amd import correlations
from oracle import ref_fp64 as R

pytestmark = pytest.mark.gpu


def field(n, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, n, n))
    d = np.exp(0.6 * g)                       # skewed: a non-zero bispectrum
    return (d / d.mean() - 1.0).astype(np.float32)


@pytest.mark.parametrize("n,L,f1,f2,nth,mas", [(16, 200.0, 2.0, 3.0, 5, "CIC"), (32, 500.0, 3.0, 5.0, 9, "CIC"),
                                               (48, 700.0, 4.0, 6.5, 6, "NONE"), (64, 1000.0, 5.0, 8.0, 15, "CIC"),
                                               (128, 1000.0, 10.0, 20.0, 8, "CIC"),
                                               (96, 900.0, 7.0, 12.0, 7, "CIC"), (24, 300.0, 2.5, 4.0, 4, "CIC")])
def test_bispec_parity(gpu, n, L, f1, f2, nth, mas):
    print(f"bispec n={n} shells={nth + 2}", flush=True)
    d = field(n, n)
    kF = 2 * np.pi / L
    theta = np.linspace(0, np.pi, nth)
    k_all, pk, th, bk, qk = (t.cpu().numpy() for t in correlations.bispec(torch.as_tensor(d, device=gpu), L,
                                                                         f1 * kF, f2 * kF, theta, mas=mas))
    o_k, o_pk, o_th, o_b, o_q, ex = R.bispec(d, L, f1 * kF, f2 * kF, theta, p=1 if mas == "CIC" else 0)
    np.testing.assert_array_equal(k_all, o_k.astype(np.float32))
    np.testing.assert_array_equal(th, o_th.astype(np.float32))
    np.testing.assert_allclose(pk, o_pk, rtol=1e-5)
    tol_b = 2e-5 * ex["scale_B"]
    assert np.all(np.abs(bk - o_b) <= tol_b), np.max(np.abs(bk - o_b) / tol_b)
    den = o_b / o_q
    tol_q = tol_b / np.abs(den) + 1e-5 * np.abs(o_q)
    assert np.all(np.abs(qk - o_q) <= tol_q), np.max(np.abs(qk - o_q) / tol_q)


def test_bispec_shell_sums_reuse_is_bit_identical(gpu):
    n, L = 64, 800.0
    d = torch.as_tensor(field(n, 7), device=gpu)
    kF = 2 * np.pi / L
    theta = np.linspace(0, np.pi, 6)
    fresh = correlations.bispec(d, L, 4 * kF, 7 * kF, theta, reuse_shell_sums=False, full=True)
    first = correlations.bispec(d, L, 4 * kF, 7 * kF, theta)      # fills or hits the cache
    again = correlations.bispec(d, L, 4 * kF, 7 * kF, theta, full=True)
    for x, y in zip(fresh[:5], again[:5]):
        assert torch.equal(x, y)
    assert torch.equal(fresh[5]["sums"], again[5]["sums"])
    for x, y in zip(first, again[:5]):
        assert torch.equal(x, y)


def test_bispec_full_size_runs_and_is_deterministic(gpu):
    n, L = 512, 1500.0
    print("bispec 512^3", flush=True)
    g = torch.Generator(device=gpu).manual_seed(3)
    d = torch.exp(0.5 * torch.randn((n, n, n), device=gpu, generator=g))
    d = d / d.mean() - 1.0
    kF = 2 * np.pi / L
    theta = np.linspace(0, np.pi, 6)
    a = correlations.bispec(d, L, 10 * kF, 20 * kF, theta)
    b = correlations.bispec(d, L, 10 * kF, 20 * kF, theta)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    pk = a[1].cpu().numpy()
    assert np.all(np.isfinite(pk)) and np.all(pk > 0)
    assert np.all(np.isfinite(a[3].cpu().numpy()))
