"""Model-side functions (jax_powspec_amd.rsd_model, src/correlations.py:274-333) on the
GPU: every function runs on CUDA float64 tensors, keeps its outputs there, and agrees with
the CPU evaluation that tests/test_rsd_model.py pins by quadrature and closed forms (to
1e-10: the GPU reductions add in another order)."""
import numpy as np
import pytest
import torch

from jax_powspec_amd import rsd_model as M

pytestmark = pytest.mark.gpu


def close(gpu_out, cpu_out):
    outs = gpu_out if isinstance(gpu_out, tuple) else (gpu_out,)
    refs = cpu_out if isinstance(cpu_out, tuple) else (cpu_out,)
    for g, c in zip(outs, refs):
        assert g.is_cuda
        np.testing.assert_allclose(g.cpu().numpy(), c.numpy(), rtol=1e-10, atol=0)


def test_kaiser_multipoles_on_gpu(gpu):
    P = torch.linspace(100, 5000, 64, dtype=torch.float64)
    close(M.kaiser_power_spectrum(P.to(gpu), 1.7, 0.8), M.kaiser_power_spectrum(P, 1.7, 0.8))
    edges = torch.linspace(0, 1, 2001, dtype=torch.float64)
    close(M.kaiser_power_spectrum_integration(P.to(gpu), edges.to(gpu), 2.0, 0.75),
          M.kaiser_power_spectrum_integration(P, edges, 2.0, 0.75))


@pytest.mark.parametrize("bugs", [False, True])
def test_fog_multipoles_on_gpu(gpu, bugs):
    k = torch.linspace(0.01, 0.4, 40, dtype=torch.float64)
    P = 2e4 * (k / 0.05) ** -1.5
    edges = torch.linspace(0, 1, 4001, dtype=torch.float64)
    close(M.kaiser_fog_power_spectrum_integrate(P.to(gpu), k.to(gpu), edges.to(gpu), 1.5, 0.7, 3.0,
                                                reference_bugs=bugs),
          M.kaiser_fog_power_spectrum_integrate(P, k, edges, 1.5, 0.7, 3.0, reference_bugs=bugs))


def test_variances_on_gpu(gpu):
    k = torch.linspace(0.0, 0.3, 61, dtype=torch.float64)   # k = 0: the nansum skip
    pk = 2e4 * torch.exp(-k / 0.1)
    var = M.estimate_pk_variance(k[1:], pk[1:], 1000.0, 300.0, 0.005)
    close(M.estimate_pk_variance(k[1:].to(gpu), pk[1:].to(gpu), 1000.0, 300.0, 0.005), var)
    s = torch.linspace(5.0, 150.0, 30, dtype=torch.float64)
    v = torch.cat([torch.zeros(1, dtype=torch.float64), var])
    cov_g = M.estimate_xi_covariance(s.to(gpu), k.to(gpu), pk.to(gpu), 2.0, v.to(gpu), 0.005)
    cov_c = M.estimate_xi_covariance(s, k, pk, 2.0, v, 0.005)
    close(cov_g, cov_c)
    assert torch.allclose(cov_g, cov_g.T)
