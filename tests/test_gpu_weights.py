"""Weighted assignment at the edges of the 64-bit fixed-point deposit (GPU, through the C ABI).

Every deposit is rint(w W S) with S = 2^30 / 2^e and max|w| < 2^e, converted through int32
(corner weights W <= 1).  These cases put max|w| exactly on a power of two, mix signs,
scale weights by 1e6 and 1e-6, and check the mesh (sorted and direct tiles) and the fused
estimate against the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from jax_powspec_amd import correlations, mas
from oracle import ref_fp64 as R
from test_gpu_parity import pk_check

pytestmark = pytest.mark.gpu
F32 = np.float32


def weights(kind, rng, npart):
    if kind == "signed":
        return rng.uniform(-3.0, 5.0, npart).astype(F32)
    if kind == "pow2_max":   # max|w| = 4 exactly: e = 3, S = 2^27
        w = rng.uniform(0.0, 4.0, npart).astype(F32)
        w[:7] = 4.0
        w[7:14] = -4.0
        return w
    if kind == "large":
        return (rng.uniform(0.5, 1.5, npart) * 1e6).astype(F32)
    if kind == "tiny":
        return (rng.uniform(0.5, 1.5, npart) * 1e-6).astype(F32)
    raise ValueError(kind)


def positions(rng, npart, box):
    p = rng.random((npart, 3)) * box
    c = rng.random((8, 3)) * box
    k = npart // 3
    p[:k] = (c[rng.integers(0, 8, k)] + rng.normal(0, box / 50, (k, 3))) % box
    return p.astype(F32)


@pytest.mark.parametrize("stencil", ["CIC", "TSC"])
@pytest.mark.parametrize("kind", ["signed", "pow2_max", "large", "tiny"])
@pytest.mark.parametrize("tile", [None, "direct"])
def test_weighted_mesh(gpu, stencil, kind, tile):
    rng = np.random.default_rng(hash((stencil, kind)) % 2**32)
    box, n, npart = 200.0, 32, 30000
    p = positions(rng, npart, box)
    w = weights(kind, rng, npart)
    t = torch.as_tensor(p, device=gpu)
    got = mas.assign(t[:, 0], t[:, 1], t[:, 2], torch.as_tensor(w, device=gpu), box_size=box, n_bins=n, mas=stencil,
                     tile=tile).cpu().numpy().astype(np.float64)
    want = (R.cic_mesh if stencil == "CIC" else R.tsc_mesh)(*p.T, w, 0, 0, 0, box, n)
    scale = np.abs(w.astype(np.float64)).max() * 30.0   # a busy cell's mass
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6 * scale)
    # the fixed-point sums are order-independent: the same catalogue shuffled gives the same bits
    perm = rng.permutation(npart)
    tp = torch.as_tensor(p[perm], device=gpu)
    again = mas.assign(tp[:, 0], tp[:, 1], tp[:, 2], torch.as_tensor(w[perm], device=gpu), box_size=box, n_bins=n,
                       mas=stencil, tile=tile).cpu().numpy().astype(np.float64)
    if tile is None:
        np.testing.assert_array_equal(again, got)


@pytest.mark.parametrize("kind", ["signed", "pow2_max", "large"])
def test_weighted_estimate(gpu, kind):
    rng = np.random.default_rng(11 + len(kind))
    box, n, npart = 500.0, 32, 40000
    p = positions(rng, npart, box)
    w = weights(kind, rng, npart)
    if kind == "signed":
        w = np.abs(w) + 0.1   # the contrast needs a positive mean density
    t = torch.as_tensor(p, device=gpu)
    kF = 2 * np.pi / box
    edges = kF * (0.5 + np.arange(n // 2 + 1))
    _, pk, nm, ex = correlations.estimate(t[:, 0], t[:, 1], t[:, 2], torch.as_tensor(w, device=gpu), box_size=box,
                                          n_bins=n, k_edges=edges, mas="CIC", full=True)
    res = R.estimate(*p.T, w, box, n, edges, stencil="CIC", p=2)
    pk_check(pk, nm, ex, res[1], res[3], box, n)
