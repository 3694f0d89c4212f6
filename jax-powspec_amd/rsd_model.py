"""Model side of the P_l(k) fits -- drop-in for src/correlations.py:274-333.

Small array functions, on whatever device their inputs live (torch tensors; NumPy arrays
and Python scalars are accepted and computed in float64 on the CPU):

* ``kaiser_power_spectrum``               <- :287-297, Eq. 2.1 of arXiv:1610.07785;
* ``kaiser_power_spectrum_integration``   <- :274-285, midpoint sums over mu bins;
* ``kaiser_fog_power_spectrum_integrate`` <- :299-310, Kaiser with a Gaussian
  finger-of-God damping exp(-k^2 mu^2 sigma^2);
* ``estimate_pk_variance``                <- :312-317, Eq. 17 of arXiv:2109.15236;
* ``estimate_xi_covariance``              <- :318-331.

Two defects of the reference's ``kaiser_fog_power_spectrum_integrate`` are corrected by
default: its l = 4 weight uses 35 mu^4 - 30 mu^2 + 2 instead of Legendre's + 3 (:308), and
its sums are not multiplied by d mu (the Kaiser integration two functions above is).
``reference_bugs=True`` reproduces the reference's numbers exactly.
"""
from __future__ import annotations

import math

import numpy as np
import torch

__all__ = ["kaiser_power_spectrum", "kaiser_power_spectrum_integration", "kaiser_fog_power_spectrum_integrate",
           "estimate_pk_variance", "estimate_xi_covariance"]


def _t(x, like=None):
    if isinstance(x, torch.Tensor):
        return x
    dev = like.device if isinstance(like, torch.Tensor) else None
    dt = like.dtype if isinstance(like, torch.Tensor) and like.is_floating_point() else torch.float64
    return torch.as_tensor(np.asarray(x, dtype=np.float64), dtype=dt, device=dev)


def kaiser_power_spectrum(power_spectrum, bias, growth_rate):
    """Linear-theory multipoles (P0, P2, P4) of a biased tracer (src/correlations.py:287-297)."""
    P = _t(power_spectrum)
    b, f = float(bias), float(growth_rate)
    beta = f / b
    mono = P * (b * b / 15.0) * (15.0 + 10.0 * beta + 3.0 * beta * beta)
    quad = P * (4.0 / 21.0) * f * (7.0 * b + 3.0 * f)
    hexa = P * (8.0 / 35.0) * f * f
    return mono, quad, hexa


def _mu_mid(mu_edges, like):
    e = _t(mu_edges, like)
    mu = 0.5 * (e[:-1] + e[1:])
    return mu, mu[1] - mu[0]


def kaiser_power_spectrum_integration(power_spectrum, mu_edges, bias, growth_rate):
    """(P0, P2, P4) from midpoint sums of P(k) (b + f mu^2)^2 over mu bins
    (src/correlations.py:274-285); mu_edges equally spaced."""
    P = _t(power_spectrum)
    mu, dmu = _mu_mid(mu_edges, P)
    K = (float(bias) + float(growth_rate) * mu * mu) ** 2
    m = P[:, None] * K[None, :]
    quad = (m * (5.0 * 0.5 * (3.0 * mu * mu - 1.0))[None, :]).sum(dim=1) * dmu
    hexa = (m * (9.0 * 0.125 * (35.0 * mu ** 4 - 30.0 * mu * mu + 3.0))[None, :]).sum(dim=1) * dmu
    mono = m.sum(dim=1) * dmu
    return mono, quad, hexa


def kaiser_fog_power_spectrum_integrate(power_spectrum, k, mu_edges, bias, growth_rate, sigma,
                                        reference_bugs: bool = False):
    """Kaiser multipoles with a Gaussian finger-of-God factor exp(-k^2 mu^2 sigma^2)
    (src/correlations.py:299-310).  reference_bugs=True keeps the reference's l = 4 constant
    (+2) and its missing d mu."""
    P = _t(power_spectrum)
    kk = _t(k, P)
    mu, dmu = _mu_mid(mu_edges, P)
    K = (float(bias) + float(growth_rate) * mu * mu) ** 2
    m = P[:, None] * K[None, :] * torch.exp(-(kk[:, None] ** 2) * (mu[None, :] ** 2) * float(sigma) ** 2)
    c4 = 2.0 if reference_bugs else 3.0
    w = 1.0 if reference_bugs else dmu
    quad = (m * (5.0 * 0.5 * (3.0 * mu * mu - 1.0))[None, :]).sum(dim=1) * w
    hexa = (m * (9.0 * 0.125 * (35.0 * mu ** 4 - 30.0 * mu * mu + c4))[None, :]).sum(dim=1) * w
    mono = m.sum(dim=1) * w
    return mono, quad, hexa


def estimate_pk_variance(k, pk, box_size, shot_noise, dk):
    """Gaussian variance of the monopole in linear k bins of width dk
    (src/correlations.py:312-317): (2 pi)^3 / L^3 * 2 (P + N)^2 / (4 pi k^2 dk)."""
    kk = _t(k)
    P = _t(pk, kk)
    return (2.0 * math.pi) ** 3 / float(box_size) ** 3 * (2.0 * (P + float(shot_noise)) ** 2
                                                          / (4.0 * math.pi * kk * kk * float(dk)))


def estimate_xi_covariance(s, k, pk, a_damp, pk_variance, dk):
    """Covariance of xi_0(s) from the P(k) variance (src/correlations.py:318-331):
    sum_k dk^2 k^4 var(k) exp(-2 k^2 a^2) / (2 pi^2)^2 j0(k s_a) j0(k s_b), NaN terms
    (k s = 0) skipped as the reference's nansum."""
    kk = _t(k)
    ss = _t(s, kk)
    var = _t(pk_variance, kk)
    ks = kk[:, None] * ss[None, :]
    damp = torch.exp(-kk * kk * float(a_damp) ** 2)
    j0 = torch.sin(ks) / ks
    factor = float(dk) ** 2 * kk ** 4 * var * damp * damp / (2.0 * math.pi ** 2) ** 2
    terms = factor[:, None, None] * j0[:, :, None] * j0[:, None, :]
    return torch.nansum(terms, dim=0)
