"""GPU parity of the configuration-space multipoles (pk_xi_mesh through the C ABI) against
the FP64 oracle (oracle/ref_fp64.py: xi_vec, xi_vec_fundamental).

Bars: bin populations exact (Nmodes, including inf / 0 in empty bins); NaN where the
oracle has NaN; |xi_l - xi_l^oracle| <= 1e-5 A_lb + 2e-6 (2l+1) g0 / n^3, where A_lb is the
same-bin absolute scale (2l+1) sum |xi_cell L_l| / (N n^3) and g0 = xi_grid(0), the largest
value of the separation grid: the second term is float32 FFT round-off, which is set by the
whole grid rather than by a bin's own values.
"""
import numpy as np
import pytest
import torch

from jax_powspec_amd import correlations
from oracle import ref_fp64 as R

pytestmark = pytest.mark.gpu


def field(n, seed, mean=0.0):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n)
    kz = np.fft.rfftfreq(n)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    f = np.fft.irfftn(np.fft.rfftn(w) * np.exp(-(kk / 0.15) ** 2) / np.maximum(kk, 1.0 / n), s=(n, n, n))
    f[..., :] += 0.3 * np.sin(2 * np.pi * np.arange(n) / n)          # some anisotropy along z
    return (mean + f / f.std()).astype(np.float32)


def check(gpu_out, ora, n, d):
    r, xi, nm = (t.cpu().numpy() for t in gpu_out)
    o_r, o_xi, o_N, ex = ora
    np.testing.assert_array_equal(nm.astype(np.float64), np.asarray(o_N, np.float64).astype(np.float32))
    assert np.array_equal(np.isnan(xi), np.isnan(o_xi))
    g0 = R.xi_grid(d)[0, 0, 0]
    tol = 1e-5 * np.nan_to_num(ex["scale"]) + 2e-6 * np.array([1.0, 5.0, 9.0]) * g0 / n ** 3
    ok = ~np.isnan(o_xi)
    err = np.abs(xi.astype(np.float64) - o_xi)
    assert np.all(err[ok] <= tol[ok]), (np.max((err / tol)[ok]))
    np.testing.assert_allclose(r, o_r.astype(np.float32), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("n,box,edges", [
    (8, 100.0, np.arange(1e-3, 60, 7.0)), (12, 300.0, np.linspace(0.0, 150, 9)),
    (16, 50.0, np.array([0.5, 3.125, 9.375, 40.0])), (32, 1000.0, np.arange(1e-3, 200, 5.0)),
    (64, 1000.0, np.arange(1e-3, 200, 5.0)), (96, 1500.0, np.arange(0.0, 900.0, 2.5)),
    (128, 1000.0, np.arange(1e-3, 200, 5.0)), (256, 1000.0, np.arange(1e-3, 200, 1.0))])
def test_xi_vec_parity(gpu, n, box, edges):
    print(f"xi_vec n={n} nb={edges.size - 1}", flush=True)
    d = field(n, n)
    out = correlations.xi_vec(torch.as_tensor(d, device=gpu), box, edges)
    check(out, R.xi_vec(d, box, edges), n, d)


@pytest.mark.parametrize("n", [8, 12, 16, 48, 64, 128, 192, 256])
def test_xi_fundamental_parity(gpu, n):
    print(f"xi_vec_fundamental n={n}", flush=True)
    box = 700.0
    d = field(n, 100 + n)
    out = correlations.xi_vec_fundamental(torch.as_tensor(d, device=gpu), box)
    check(out, R.xi_vec_fundamental(d, box), n, d)


def test_xi_full_size_sum_rule_and_determinism(gpu):
    """512^3: every separation in one bin gives sum_s xi_grid(s) = (sum delta)^2; two runs
    are bit-identical (fixed-point bins)."""
    n, box = 512, 1000.0
    print("xi 512^3", flush=True)
    g = torch.Generator(device=gpu).manual_seed(5)
    d = 1.0 + 0.2 * torch.randn((n, n, n), device=gpu, generator=g)
    edges = np.array([-1.0, 2.0 * box])
    _, xi, nm, ex = correlations.xi_vec(d, box, edges, mas="NONE", full=True)
    total = ex["sums"][0, 0].item()
    want = d.double().sum().item() ** 2
    assert abs(total / want - 1) < 1e-5, (total, want)
    assert int(nm[0].item()) == n ** 3
    edges = np.arange(1e-3, 200, 2.0)
    a = correlations.xi_vec(d - 1.0, box, edges)
    b = correlations.xi_vec(d - 1.0, box, edges)
    assert torch.equal(a[1], b[1])
