"""pk_powspec_mesh_noalloc (the allocation-free variant of pk_powspec_mesh) through the C ABI.

It must return what pk_powspec_mesh returns on the same inputs: identical mode counts,
FP64 sums within 1e-12 of the same-bin scale (both add FP64 partials with LDS atomics whose
order follows wave scheduling), float32 P_l the rounding of those sums.  Two calls in a
row, the second on a smaller mesh with fewer bins, so the second reuses scratch the first
left filled.  Requests above the fixed scratch are rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from jax_powspec_amd import _device, _lib, binning

pytestmark = pytest.mark.gpu


def _mesh(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n, n)).astype(np.float32)
    # a little large-scale structure so P_2, P_4 are not pure noise
    x = np.arange(n)
    a += np.cos(2 * np.pi * 3 * x / n).astype(np.float32)[None, None, :]
    return a


def _outputs(nb, dev):
    return dict(sums=torch.full((nb, 3), np.nan, dtype=torch.float64, device=dev),
                counts=torch.full((nb,), -1, dtype=torch.int64, device=dev),
                pk=torch.full((nb, 3), np.nan, dtype=torch.float32, device=dev),
                nmodes=torch.full((nb,), -1.0, dtype=torch.float32, device=dev),
                kmean=torch.full((nb,), np.nan, dtype=torch.float64, device=dev))


def _run(noalloc, mesh, n, box, edges, dev):
    lib = _lib.load()
    plan = _device.plan(n, dev)
    tab = binning.TABLES.get(dev, n, binning.grid_edges(edges, box))
    nb = tab.nb
    o = _outputs(nb, dev)
    spec = torch.empty((n // 2 + 1, n, n, 2), dtype=torch.float32, device=dev)
    common = (plan.handle, _device.ptr(mesh), None, _device.ptr(tab.thr), _device.ptr(tab.jst),
              _device.ptr(tab.modes), _device.ptr(tab.ksum), nb, 2, _lib.FLAG_KMEAN, float(box), 0.0,
              _device.ptr(spec), _device.ptr(o["sums"]), _device.ptr(o["counts"]), _device.ptr(o["pk"]),
              _device.ptr(o["nmodes"]), _device.ptr(o["kmean"]))
    stream = _device.stream_ptr(dev)
    if noalloc:
        _lib.check("pk_powspec_mesh_noalloc", lib.pk_powspec_mesh_noalloc(*common, stream))
    else:
        nbytes = C.c_size_t()
        _lib.check("pk_powspec_workspace_size", lib.pk_powspec_workspace_size(plan.handle, nb, C.byref(nbytes)))
        ws = _device.workspace(nbytes.value, dev)
        _lib.check("pk_powspec_mesh", lib.pk_powspec_mesh(*common, _device.ptr(ws), nbytes.value, stream))
    torch.cuda.synchronize(dev)
    return {k: v.cpu().numpy() for k, v in o.items()}


def _same(a, b):
    np.testing.assert_array_equal(a["counts"], b["counts"])
    np.testing.assert_array_equal(a["nmodes"], b["nmodes"])
    scale = np.abs(b["sums"][:, :1]) + 1e-300
    assert np.all(np.abs(a["sums"] - b["sums"]) <= 1e-12 * scale), np.max(np.abs(a["sums"] - b["sums"]) / scale)
    ok = b["counts"] > 0
    np.testing.assert_allclose(a["pk"][ok], b["pk"][ok], rtol=2e-7, atol=0)
    assert np.all(np.isnan(a["pk"][~ok])) and np.all(np.isnan(b["pk"][~ok]))
    np.testing.assert_allclose(a["kmean"][ok], b["kmean"][ok], rtol=1e-12)


def test_noalloc_matches_workspace_entry_point(gpu):
    box = 1000.0
    kF = 2 * np.pi / box
    cases = [(128, kF * (0.5 + np.arange(65))), (64, kF * (0.5 + np.arange(0, 33, 2)))]   # then smaller n, nb
    for n, edges in cases:
        mesh = torch.as_tensor(_mesh(n, n), device=gpu)
        print(f"[noalloc] n={n} nb={len(edges) - 1}", flush=True)
        got = _run(True, mesh, n, box, edges, gpu)
        want = _run(False, mesh, n, box, edges, gpu)
        _same(got, want)
        assert got["counts"].sum() > 0


def test_noalloc_rejects_sizes_above_the_scratch(gpu):
    n, box = 512, 1000.0
    kF = 2 * np.pi / box
    edges = np.linspace(0.5 * kF, 256 * kF, 4097)   # 4096 bins: partials far above the fixed scratch
    plan = _device.plan(n, gpu)
    nbytes = C.c_size_t()
    lib = _lib.load()
    tab = binning.TABLES.get(gpu, n, binning.grid_edges(edges, box))
    _lib.check("pk_powspec_workspace_size", lib.pk_powspec_workspace_size(plan.handle, tab.nb, C.byref(nbytes)))
    assert nbytes.value > _lib.NOALLOC_SCRATCH_BYTES
    mesh = torch.zeros((n, n, n), dtype=torch.float32, device=gpu)
    o = _outputs(tab.nb, gpu)
    spec = torch.empty((n // 2 + 1, n, n, 2), dtype=torch.float32, device=gpu)
    code = lib.pk_powspec_mesh_noalloc(plan.handle, _device.ptr(mesh), None, _device.ptr(tab.thr),
                                       _device.ptr(tab.jst), _device.ptr(tab.modes), _device.ptr(tab.ksum), tab.nb,
                                       2, 0, box, 0.0, _device.ptr(spec), _device.ptr(o["sums"]),
                                       _device.ptr(o["counts"]), _device.ptr(o["pk"]), _device.ptr(o["nmodes"]),
                                       None, _device.stream_ptr(gpu))
    assert code == _lib.PK_E_WORKSPACE
    assert b"scratch" in lib.pk_last_error()
    torch.cuda.synchronize(gpu)
    assert torch.isnan(o["sums"]).all()   # nothing was launched
