"""Model-side functions (jax_powspec_amd.rsd_model, src/correlations.py:274-333), pinned by
the closed forms and quadratures they approximate rather than by their own formulas."""
import math

import numpy as np
import pytest
import torch
from scipy import integrate, special

from jax_powspec_amd import rsd_model as M


def legendre(l, mu):
    return special.eval_legendre(l, mu)


def test_kaiser_closed_form_matches_quadrature():
    P = torch.tensor([1000.0, 2500.0], dtype=torch.float64)
    b, f = 1.7, 0.8
    mono, quad, hexa = M.kaiser_power_spectrum(P, b, f)
    for l, got in ((0, mono), (2, quad), (4, hexa)):
        want = (2 * l + 1) * integrate.quad(lambda mu: (b + f * mu * mu) ** 2 * legendre(l, mu), 0, 1)[0]
        np.testing.assert_allclose(got.numpy(), P.numpy() * want, rtol=1e-12)


def test_kaiser_integration_converges_to_closed_form():
    P = torch.linspace(100, 5000, 7, dtype=torch.float64)
    b, f = 2.0, 0.75
    edges = np.linspace(0, 1, 4001)
    got = M.kaiser_power_spectrum_integration(P, edges, b, f)
    want = M.kaiser_power_spectrum(P, b, f)
    # midpoint rule: error h^2 / 24 |f'(1) - f'(0)| with h = 1/4000, |f'| < 1e3 for l = 4
    for g, w in zip(got, want):
        np.testing.assert_allclose(g.numpy(), w.numpy(), rtol=5e-5)


@pytest.mark.parametrize("sigma", [0.0, 3.0])
def test_fog_against_quadrature(sigma):
    k = np.array([0.02, 0.1, 0.25])
    P = torch.as_tensor(np.array([8000.0, 3000.0, 900.0]))
    b, f = 1.5, 0.7
    edges = np.linspace(0, 1, 8001)
    mono, quad, hexa = M.kaiser_fog_power_spectrum_integrate(P, k, edges, b, f, sigma)
    for l, got in ((0, mono), (2, quad), (4, hexa)):
        for i in range(k.size):
            want = (2 * l + 1) * float(P[i]) * integrate.quad(
                lambda mu: (b + f * mu * mu) ** 2 * math.exp(-(k[i] * mu * sigma) ** 2) * legendre(l, mu), 0, 1)[0]
            assert abs(float(got[i]) - want) <= 1e-6 * abs(float(P[i])) * (2 * l + 1), (l, i, float(got[i]), want)


def test_fog_reference_bugs_flag():
    k = np.array([0.05])
    P = torch.tensor([1000.0], dtype=torch.float64)
    edges = np.linspace(0, 1, 11)
    fixed = M.kaiser_fog_power_spectrum_integrate(P, k, edges, 1.0, 0.5, 2.0)
    ref = M.kaiser_fog_power_spectrum_integrate(P, k, edges, 1.0, 0.5, 2.0, reference_bugs=True)
    dmu = 0.1
    np.testing.assert_allclose(ref[0].numpy() * dmu, fixed[0].numpy(), rtol=1e-12)
    np.testing.assert_allclose(ref[1].numpy() * dmu, fixed[1].numpy(), rtol=1e-12)
    # hexadecapole differs by 9 * 0.125 * (2 - 3) * sum(m) = -(9/8) mono_sum
    np.testing.assert_allclose((ref[2] - (-9 / 8) * ref[0]).numpy() * dmu, fixed[2].numpy(), rtol=1e-9)


def test_pk_variance_is_gaussian_mode_count():
    # var = 2 (P + N)^2 / N_modes with N_modes = 4 pi k^2 dk L^3 / (2 pi)^3
    k = np.array([0.05, 0.1])
    pk, L, N, dk = np.array([5000.0, 2000.0]), 1000.0, 300.0, 0.005
    var = M.estimate_pk_variance(k, pk, L, N, dk).numpy()
    nm = 4 * np.pi * k ** 2 * dk * L ** 3 / (2 * np.pi) ** 3
    np.testing.assert_allclose(var, 2 * (pk + N) ** 2 / nm, rtol=1e-12)


def test_xi_covariance_sum():
    k = np.linspace(0.0, 0.3, 31)        # k = 0 gives 0/0 terms, skipped like nansum
    s = np.array([10.0, 40.0, 90.0])
    var = np.full(k.size, 2e6)
    cov = M.estimate_xi_covariance(s, k, np.zeros_like(k), 2.0, var, 0.01).numpy()
    want = np.zeros((3, 3))
    for kk, v in zip(k[1:], var[1:]):
        j = np.sin(kk * s) / (kk * s)
        want += 0.01 ** 2 * kk ** 4 * v * math.exp(-2 * kk * kk * 4.0) / (2 * math.pi ** 2) ** 2 * np.outer(j, j)
    np.testing.assert_allclose(cov, want, rtol=1e-12)
    assert np.allclose(cov, cov.T)
