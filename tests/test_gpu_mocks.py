"""GPU parity of the seeded mock generator (libpowspec_amd.so through the C ABI) against
the CPU oracle (oracle/mocks.py).

Bars:
  * structure exact: delta(0) = 0, self-conjugate modes real, (k, -k) pairs on the kz = 0
    and kz = n/2 planes exact conjugates;
  * half-spectrum modes within float32 rounding of the FP64 oracle (1e-5 relative);
  * irfftn within float32 FFT rounding of numpy's FP64 irfftn (1e-5 of the field's max);
  * the fused mesh path bit-identical to spectrum + irfftn;
  * populate: particle totals and per-cell counts exact, positions within 1e-6 of the box;
  * whole chain: total count within 6 sigma of density L^3; with the redshift-space shift,
    the monopole gain within 10 % and the quadrupole gain within 15 % of Kaiser's.
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from jax_powspec_amd import _device, _lib, correlations, gauss_field as GF
from oracle import mocks as M

pytestmark = pytest.mark.gpu


def table(peak=2e4, k0=0.05):
    kf = np.linspace(1e-4, 3.0, 600)
    Pkf = peak * (kf / k0) / (1.0 + (kf / k0) ** 2.5)
    return kf, Pkf


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,box,rayleigh,seed", [
    (8, 300.0, 0, 11), (12, 300.0, 1, 12), (16, 500.0, 1, 2**40 + 5), (32, 800.0, 0, 3), (64, 1000.0, 1, 77),
    (96, 1500.0, 1, 123456789)])
def test_gaussian_field_parity(gpu, n, box, rayleigh, seed):
    say(f"gaussian_field n={n} rayleigh={rayleigh} seed={seed}")
    kf, Pkf = table()
    d = GF.gaussian_field(n, kf, Pkf, rayleigh, seed, box).cpu().numpy()
    assert d.shape == (n, n, n // 2 + 1) and d.dtype == np.complex64
    o = M.gaussian_field(n, kf, Pkf, rayleigh, seed, box)
    err = np.abs(d.astype(np.complex128) - o)
    assert np.all(err <= 1e-5 * np.abs(o) + 1e-7 * np.abs(o).max()), err.max()
    # exact structure
    h = n // 2
    assert d[0, 0, 0] == 0
    for kz in (0, h):
        p = d[:, :, kz]
        mirror = p[(-np.arange(n)) % n][:, (-np.arange(n)) % n]
        assert np.array_equal(p, np.conj(mirror))
        for a in (0, h):
            for b in (0, h):
                assert p[a, b].imag == 0


@pytest.mark.parametrize("n", [8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256])
def test_irfftn_parity(gpu, n):
    say(f"irfftn n={n}")
    rng = np.random.default_rng(n)
    # an arbitrary (non-Hermitian) half spectrum: numpy's handling of the kz = 0 and n/2
    # planes (imaginary parts dropped after the x and y inverses) must be reproduced
    spec = (rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))).astype(np.complex64)
    t = torch.as_tensor(spec, device=gpu)
    before = t.clone()
    g = GF.irfftn(t).cpu().numpy().astype(np.float64)
    o = np.fft.irfftn(spec.astype(np.complex128), s=(n, n, n), axes=(0, 1, 2))
    scale = np.abs(o).max()
    assert np.abs(g - o).max() <= 1e-5 * scale, (np.abs(g - o).max(), scale)
    assert torch.equal(t, before)     # the input is only read


@pytest.mark.parametrize("n,rayleigh,seed", [(16, 1, 5), (24, 1, 6), (64, 0, 9), (96, 1, 10), (128, 1, 2**33 + 1),
                                             (192, 0, 12)])
def test_gaussian_mesh_fused(gpu, n, rayleigh, seed):
    say(f"gaussian_mesh n={n}")
    box = 1000.0
    kf, Pkf = table()
    spec = GF.gaussian_field(n, kf, Pkf, rayleigh, seed, box)
    two_step = GF.irfftn(spec)
    fused = GF.gaussian_mesh(n, kf, Pkf, rayleigh, seed, box)
    assert torch.equal(two_step, fused)
    o = M.irfftn(M.gaussian_field(n, kf, Pkf, rayleigh, seed, box))
    g = fused.cpu().numpy().astype(np.float64)
    assert np.abs(g - o).max() <= 1e-5 * np.abs(o).max()
    psi = GF.gaussian_mesh(n, kf, Pkf, rayleigh, seed, box, component="psi_z").cpu().numpy().astype(np.float64)
    op = M.irfftn(M.psi_z_spectrum(M.gaussian_field(n, kf, Pkf, rayleigh, seed, box), box))
    assert np.abs(psi - op).max() <= 1e-5 * np.abs(op).max()


@pytest.mark.parametrize("n", [256, 384, 512])
def test_gaussian_mesh_round_trip_full_size(gpu, n):
    """rfftn(gaussian_mesh) = gaussian_field at full size (the field is exactly Hermitian)."""
    say(f"round trip n={n}")
    box = 1000.0
    kf, Pkf = table()
    g = GF.gaussian_mesh(n, kf, Pkf, 1, 31, box)
    back = correlations.rfftn(g)
    spec = GF.gaussian_field(n, kf, Pkf, 1, 31, box)
    diff = (back - spec).abs()
    ref = spec.abs()
    rel = (diff.square().sum() / ref.square().sum()).sqrt().item()
    assert rel < 2e-6, rel
    assert (diff.max() / ref.max()).item() < 2e-5
    del g, back, spec, diff, ref
    torch.cuda.empty_cache()


def _field(rng, n, kind):
    if kind == "lognormal":
        return np.exp(0.8 * rng.standard_normal((n, n, n))).astype(np.float32)
    if kind == "sparse":   # zeros, negatives (rate 0) and a few dense cells (PTRS regime)
        f = rng.random((n, n, n)).astype(np.float32)
        f[rng.random((n, n, n)) < 0.3] = 0.0
        f[rng.random((n, n, n)) < 0.02] = -1.0
        idx = rng.integers(0, n, (40, 3))
        f[idx[:, 0], idx[:, 1], idx[:, 2]] = rng.random(40).astype(np.float32) * 400 + 50
        return f
    return rng.standard_normal((n, n, n)).astype(np.float32)     # Gaussian (bias != 0)


@pytest.mark.parametrize("n,kind,density,bias,seed", [
    (8, "lognormal", 0.5, 0.0, 1), (12, "lognormal", 0.2, 0.0, 2), (16, "sparse", 0.3, 0.0, 3),
    (24, "gauss", 0.05, 1.3, 4), (32, "lognormal", 0.002, 0.0, 2**36 + 7), (48, "sparse", 0.01, 0.0, 8),
    (64, "gauss", 0.004, 1.1, 100), (96, "lognormal", 0.0005, 0.0, 17), (7, "lognormal", 1.0, 0.0, 21)])
def test_populate_parity(gpu, n, kind, density, bias, seed):
    box = 250.0 + 10 * n
    say(f"populate n={n} kind={kind} density={density} bias={bias}")
    rng = np.random.default_rng(seed % 1000)
    f = _field(rng, n, kind)
    pos = GF.populate_field(torch.as_tensor(f, device=gpu), n, box, density, seed, bias=bias).cpu().numpy()
    opos, counts = M.populate(f, n, box, density, seed, bias=bias)
    assert pos.shape == opos.shape, (pos.shape, opos.shape)
    assert pos.dtype == np.float32
    assert np.all(pos >= 0) and np.all(pos < box)
    assert np.abs(pos.astype(np.float64) - opos).max() <= 1e-6 * box
    # particles come in C order of their cells, so equal arrays mean equal per-cell counts
    assert counts.sum() == pos.shape[0]


def test_populate_rsd_parity(gpu):
    n, box, density, seed = 32, 600.0, 0.01, 5
    rng = np.random.default_rng(1)
    f = rng.standard_normal((n, n, n)).astype(np.float32)
    psi = (rng.standard_normal((n, n, n)) * 30).astype(np.float32)    # shifts beyond a box too
    say("populate rsd")
    pos = GF.populate_field(torch.as_tensor(f, device=gpu), n, box, density, seed, bias=1.0,
                            psi_z=torch.as_tensor(psi, device=gpu), growth_rate=0.7).cpu().numpy()
    opos, _ = M.populate(f, n, box, density, seed, bias=1.0, psi_z=psi, growth_rate=0.7)
    assert pos.shape == opos.shape
    d = np.abs(pos.astype(np.float64) - opos)
    d = np.minimum(d, box - d)    # a value at the wrap point may land on either side
    assert d.max() <= 1e-6 * box


def test_populate_empty_and_capacity(gpu):
    n, box = 16, 100.0
    f = torch.ones((n, n, n), device=gpu)
    assert GF.populate_field(f, n, box, 0.0, 3).shape == (0, 3)
    # capacity below the total: exactly `capacity` records written, nothing past them
    lib = _lib.load()
    nbytes = C.c_size_t()
    _lib.check("ws", lib.pk_populate_workspace_size(n, C.byref(nbytes)))
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device=gpu)
    total = torch.zeros(1, dtype=torch.int64, device=gpu)
    s = _device.stream_ptr(gpu)
    _lib.check("count", lib.pk_populate_count(f.data_ptr(), n, box, 0.5, 0.0, 9, ws.data_ptr(), nbytes.value,
                                              total.data_ptr(), s))
    tot = int(total.item())
    cap = tot // 2
    buf = torch.full((tot, 3), -7.0, device=gpu)
    _lib.check("place", lib.pk_populate_place(f.data_ptr(), None, 0.0, n, box, 0.0, 9, ws.data_ptr(), nbytes.value,
                                              buf.data_ptr(), cap, s))
    full = GF.populate_field(f, n, box, 0.5, 9)
    assert full.shape[0] == tot
    assert torch.equal(buf[:cap], full[:cap])
    assert bool((buf[cap:] == -7.0).all())
    with pytest.raises(ValueError):
        lib_rc = lib.pk_populate_place(f.data_ptr(), None, 0.0, n, box, 0.0, 9, ws.data_ptr(), nbytes.value,
                                       None, -1, s)
        _lib.check("place", lib_rc)


def test_lognormal_mock_total_and_determinism(gpu):
    n, box, density, bias = 128, 1000.0, 2e-3, 1.1
    kf, Pkf = table()
    say("lognormal mock total")
    a = GF.lognormal_mock(n, kf, Pkf, box, density, bias, 42)
    b = GF.lognormal_mock(n, kf, Pkf, box, density, bias, 42)
    assert torch.equal(a, b)
    mean = density * box ** 3
    assert abs(a.shape[0] - mean) < 6 * math.sqrt(mean)
    c = GF.lognormal_mock(n, kf, Pkf, box, density, bias, 43)
    assert not torch.equal(a[:1000], c[:1000])


def test_lognormal_mock_kaiser_quadrupole(gpu):
    """Redshift-space shift z += f psi_z against the same seed without it.  Per mode
    delta_s = (1 + beta mu^2) delta_r (beta = f / b for rho = exp(b g) ~ 1 + b g), so with
    shared draws P0(f) / P0(0) -> 1 + 2 beta / 3 + beta^2 / 5 and
    (P2(f) - P2(0)) / P0(0) -> 4 beta / 3 + 4 beta^2 / 7 on large scales, with most of the
    cosmic variance cancelled.  A weakly non-linear field (b sigma_g ~ 0.25) keeps the
    lognormal's extra power small."""
    n, box, density, b, f = 128, 1000.0, 0.02, 1.5, 0.8
    kf, Pkf = table(peak=3e2)
    edges = np.linspace(0.01, 0.12, 12)
    say("kaiser")
    res = {}
    for growth in (0.0, f):
        pos = GF.lognormal_mock(n, kf, Pkf, box, density, b, 2024, growth_rate=growth)
        npart = pos.shape[0]
        k, pk, nm = correlations.estimate(pos[:, 0], pos[:, 1], pos[:, 2], box_size=box, n_bins=n, k_edges=edges,
                                          mas="CIC", shot_noise=box ** 3 / npart)
        pk = pk.cpu().numpy().astype(np.float64)
        w = nm.cpu().numpy().astype(np.float64)
        res[growth] = ((pk[:, 0] * w).sum(), (pk[:, 1] * w).sum())
    beta = f / b
    r0 = res[f][0] / res[0.0][0]
    r2 = (res[f][1] - res[0.0][1]) / res[0.0][0]
    e0 = 1 + 2 * beta / 3 + beta ** 2 / 5
    e2 = 4 * beta / 3 + 4 * beta ** 2 / 7
    say(f"P0 ratio {r0:.4f} (Kaiser {e0:.4f}), quadrupole gain {r2:.4f} (Kaiser {e2:.4f})")
    assert abs(r0 / e0 - 1) < 0.1, (r0, e0)
    assert abs(r2 / e2 - 1) < 0.15, (r2, e2)
