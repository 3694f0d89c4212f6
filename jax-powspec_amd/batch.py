"""Batched P_l(k) over many catalogues: one launch per stage for every catalogue of a mesh side.

No reference counterpart: the reference estimates one catalogue per call (e.g. its
covariance script loops over mocks, tests/estimate_pk_covariance.py).  Here every
catalogue of a batch is a job in a device-resident job table; each stage of the path
(sort, fused deposit + z pass, y pass, x pass + binning, finalize) is a single launch
over all jobs (blockIdx.y = job), optionally captured once into a HIP graph and replayed.

Each job's result equals ``correlations.estimate`` on the same catalogue with the same
tile: the kernels are the same, only the grid's y dimension selects the job.

    catalogues = [dict(x=..., y=..., z=..., box_size=L, n_bins=n, mas="CIC", shot_noise=L**3/N), ...]
    results = powspec_batched(catalogues)          # [(k3D, Pk3D, Nmodes3D), ...] in input order

``BatchEstimator`` keeps a batch resident (plan, workspace slots, job table, graph) for
repeated runs over refilled particle buffers; ``autotune`` picks the fused deposit's
row-bundle height and the launch knobs (counting chunk, y- and x-pass grids) for a mesh
side by timing every combination on a sample batch.
"""
from __future__ import annotations

import ctypes as C
from collections import defaultdict

import numpy as np
import torch

from . import _device, _lib, binning
from .correlations import _window_p

__all__ = ["BatchEstimator", "powspec_batched", "autotune", "autotune_tile", "tile_candidates", "run_concurrent",
           "TUNED_TILES", "TUNED_KNOBS"]

# n -> (1, ty, 0) chosen by autotune_tile (per process)
TUNED_TILES: dict = {}
# n -> (count_rounds, ypass_pct, xpass_pct) chosen by autotune (per process)
TUNED_KNOBS: dict = {}
DEFAULT_KNOBS = (4, 100, 100)
COUNT_ROUNDS = (1, 2, 4, 8)
PASS_PCTS = (50, 100)


def tile_candidates(n: int):
    """Row-bundle heights with a compiled fused kernel: the default (largest divisor of n
    that is <= 32 whose 64-bit fixed-point rows fit 128 KB of LDS: 32 up to n = 512, 16 at
    768 and 1024) and half of it for n >= 64."""
    d32 = n if n <= 32 else (32 if n % 32 == 0 else (24 if n % 24 == 0 else 16))
    default = d32 if d32 * n * 8 <= 131072 else 16
    out = [default]
    if n >= 64 and default % 2 == 0 and n % (default // 2) == 0:
        out.append(default // 2)
    return [(1, ty, 0) for ty in out]


class _Job:
    """Host-side state of one catalogue in a batch (tensors kept alive while queued)."""

    def __init__(self, cat: dict, n: int, device):
        self.n = n
        mas = str(cat.get("mas", "CIC_REF")).upper()
        if mas not in _lib.STENCIL:
            raise ValueError(f"unknown mass-assignment scheme {mas!r}")
        self.mas = mas
        self.box = float(cat["box_size"])
        xs, self.sx = _device.as_f32_vector(cat["x"], device, "x")
        ys, self.sy = _device.as_f32_vector(cat["y"], device, "y")
        zs, self.sz = _device.as_f32_vector(cat["z"], device, "z")
        self.np = xs.numel()
        if ys.numel() != self.np or zs.numel() != self.np:
            raise ValueError("x, y, z must have the same length")
        self.xs, self.ys, self.zs = xs, ys, zs
        self.ws, self.sw = (None, 1)
        if cat.get("w") is not None:
            self.ws, self.sw = _device.as_f32_vector(cat["w"], device, "w")
        k_edges = cat.get("k_edges")
        self.fundamental = k_edges is None
        self.edges = (binning.fundamental_grid_edges(n) if self.fundamental
                      else binning.grid_edges(k_edges, self.box))
        self.tables = binning.TABLES.get(device, n, self.edges)
        self.nb = self.tables.nb
        window = cat.get("window")
        self.p = _window_p(window if window is not None else ("TSC" if mas == "TSC" else "CIC"))
        self.flags = _lib.FLAG_KMEAN if (self.fundamental or cat.get("full")) else 0
        self.shot = float(cat.get("shot_noise", 0.0))
        self.wrap = _lib.WRAP_VEC_TRUE if cat.get("wrap", True) else _lib.WRAP_VEC_FALSE
        self.mins = tuple(_device.as_f32_scalar(cat.get(k, 0.0)) for k in ("xmin", "ymin", "zmin"))


class BatchEstimator:
    """A resident batch of up to ``max_jobs`` catalogues of mesh side ``n``.

    ``set(catalogues)`` binds particle tensors and (re)uploads the job table on the current
    stream; ``run()`` launches the batch (graph replay after the first run of a shape);
    ``results()`` returns per-job (k3D, Pk3D, Nmodes3D) views of the batch outputs.
    Outputs are overwritten by every run.  Workspace: one fixed slot per job (spectrum,
    sort buffers and bin partials sized for ``max_particles`` and ``max_bins``).
    """

    def __init__(self, n: int, max_jobs: int, max_particles: int, max_bins: int, *, tile=None, graph: bool = True,
                 device=None, capacity_grids: bool = False):
        self.device = _device.require_gpu(device)
        self.n = int(n)
        self.max_jobs = int(max_jobs)
        self.max_particles = int(max_particles)
        self.max_bins = int(max_bins)
        self.plan = _device.plan(self.n, self.device)
        self._tile_ptr, self._tile_keep = _lib.tile_arg(tile)
        lib = _lib.load()
        nbytes = C.c_size_t()
        with torch.cuda.device(self.device):
            _lib.check("pk_batch_workspace_size", lib.pk_batch_workspace_size(
                self.plan.handle, self.max_jobs, self.max_particles, self.max_bins, self._tile_ptr, C.byref(nbytes)))
            self.ws_bytes = nbytes.value
            self.ws = _device.workspace(self.ws_bytes, self.device)
            h = C.c_void_p()
            _lib.check("pk_batch_create", lib.pk_batch_create(
                self.plan.handle, self.max_jobs, self.max_particles, self.max_bins, self._tile_ptr,
                (_lib.BATCH_GRAPH if graph else 0) | (_lib.BATCH_CAPACITY_GRIDS if capacity_grids else 0),
                _device.ptr(self.ws), self.ws_bytes, C.byref(h)))
        self.handle = h
        t = (C.c_int * 3)()
        _lib.check("pk_batch_tile", lib.pk_batch_tile(self.handle, C.cast(t, C.c_void_p)))
        self.tile = (t[0], t[1], t[2])
        mj, mb = self.max_jobs, self.max_bins
        dev = self.device
        self.sums = torch.empty((mj, mb, 3), dtype=torch.float64, device=dev)
        self.counts = torch.empty((mj, mb), dtype=torch.int64, device=dev)
        self.pk = torch.empty((mj, mb, 3), dtype=torch.float32, device=dev)
        self.nmodes = torch.empty((mj, mb), dtype=torch.float32, device=dev)
        self.kmean = torch.empty((mj, mb), dtype=torch.float64, device=dev)
        self.stats = torch.zeros((mj, _lib.STATS_LEN), dtype=torch.float64, device=dev)
        self.jobs: list[_Job] = []
        self._table = (_lib.PkJob * self.max_jobs)()
        # graphs cannot capture the legacy default stream: the batch works on its own
        # stream, ordered after / before the caller's stream with events
        self._stream = torch.cuda.Stream(device=self.device)

    def _on_stream(self, stream, fn):
        cur = stream if stream is not None else torch.cuda.current_stream(self.device)
        side = self._stream
        side.wait_stream(cur)
        with torch.cuda.device(self.device):
            fn(side.cuda_stream)
        cur.wait_stream(side)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
                _lib.load().pk_batch_destroy(h)
            except Exception:
                pass
            self.handle = None

    def set(self, catalogues, *, stream=None):
        cats = list(catalogues)
        if len(cats) > self.max_jobs:
            raise ValueError(f"{len(cats)} catalogues exceed the batch capacity {self.max_jobs}")
        jobs = [c if isinstance(c, _Job) else _Job(c, self.n, self.device) for c in cats]
        for i, j in enumerate(jobs):
            if j.np > self.max_particles:
                raise ValueError(f"catalogue {i}: {j.np} particles exceed the batch capacity {self.max_particles}")
            if j.nb > self.max_bins:
                raise ValueError(f"catalogue {i}: {j.nb} bins exceed the batch capacity {self.max_bins}")
            q = self._table[i]
            q.x, q.sx = _device.ptr(j.xs), j.sx
            q.y, q.sy = _device.ptr(j.ys), j.sy
            q.z, q.sz = _device.ptr(j.zs), j.sz
            q.w, q.sw = _device.ptr(j.ws), j.sw
            q.np = j.np
            q.xmin, q.ymin, q.zmin = j.mins
            q.box_size = j.box
            q.stencil = _lib.STENCIL[j.mas]
            q.wrap_mode = j.wrap
            q.window_p = j.p
            q.flags = j.flags
            t = j.tables
            q.thr, q.jstart, q.modes, q.ksum = (_device.ptr(t.thr), _device.ptr(t.jst), _device.ptr(t.modes),
                                                _device.ptr(t.ksum))
            q.nb = j.nb
            q.shot_noise = j.shot
            q.sums = _device.ptr(self.sums[i])
            q.counts = _device.ptr(self.counts[i])
            q.pk = _device.ptr(self.pk[i])
            q.nmodes = _device.ptr(self.nmodes[i])
            q.kmean = _device.ptr(self.kmean[i])
            q.stats = _device.ptr(self.stats[i])
        self._on_stream(stream, lambda s: _lib.check("pk_batch_set_jobs", _lib.load().pk_batch_set_jobs(
            self.handle, len(jobs), C.cast(self._table, C.c_void_p), s)))
        self.jobs = jobs
        return self

    def run(self, *, stream=None):
        self._on_stream(stream, lambda s: _lib.check("pk_batch_run", _lib.load().pk_batch_run(self.handle, s)))
        return self

    def tune(self, count_rounds: int = 4, ypass_pct: int = 100, xpass_pct: int = 100):
        """Run-time knobs (pk_batch_set_tuning): counting rounds of one block per CU over all
        particles (applies at the next ``set``) and the y / x passes' blocks per job in
        percent of their defaults.  Drops captured graphs."""
        _lib.check("pk_batch_set_tuning", _lib.load().pk_batch_set_tuning(
            self.handle, int(count_rounds), int(ypass_pct), int(xpass_pct)))
        self.knobs = (int(count_rounds), int(ypass_pct), int(xpass_pct))
        return self

    def result(self, i: int, full: bool = False):
        j = self.jobs[i]
        nb = j.nb
        kF = 2.0 * np.pi / j.box
        if j.fundamental:
            k3D = (self.kmean[i, :nb] * kF).to(torch.float32)
        else:
            k3D = torch.as_tensor((0.5 * (j.edges[1:] + j.edges[:-1]) * kF).astype(np.float32), device=self.device)
        res = (k3D, self.pk[i, :nb], self.nmodes[i, :nb])
        if full:
            extras = {"sums": self.sums[i, :nb], "counts": self.counts[i, :nb],
                      "stats": self.stats[i]}
            if j.flags & _lib.FLAG_KMEAN:
                extras["kmean"] = self.kmean[i, :nb] * kF
            res = res + (extras,)
        return res

    def results(self, full: bool = False):
        return [self.result(i, full) for i in range(len(self.jobs))]


def run_concurrent(batches, *, stream=None):
    """Run several batches (e.g. one per mesh side) concurrently, each on its own stream,
    all ordered after the caller's stream and joined back into it."""
    if not batches:
        return
    dev = batches[0].device
    cur = stream if stream is not None else torch.cuda.current_stream(dev)
    lib = _lib.load()
    with torch.cuda.device(dev):
        for b in batches:
            b._stream.wait_stream(cur)
        for b in batches:
            _lib.check("pk_batch_run", lib.pk_batch_run(b.handle, b._stream.cuda_stream))
        for b in batches:
            cur.wait_stream(b._stream)


def _time_batch(b, reps):
    for _ in range(2):   # first run: plain launches; second: graph capture
        b.run()
    torch.cuda.synchronize(b.device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        b.run()
    ev1.record()
    ev1.synchronize()
    return ev0.elapsed_time(ev1) / reps


def autotune(n: int, catalogues, *, reps: int = 3, device=None, graph: bool = True):
    """Search the batched path's tile shape and launch knobs on ``catalogues`` (one mesh
    side): every compiled row-bundle height, times counting rounds (the sort's chunk size)
    in COUNT_ROUNDS, times the y- and x-pass grids at PASS_PCTS of their defaults.  Each
    configuration is timed as graph-replayed batch runs.

    Returns (best, timings): best = {"tile": (1, ty, 0), "count_rounds", "ypass_pct",
    "xpass_pct", "ms"} and timings = {(ty, rounds, ypct, xpct): ms}; the choice is cached in
    TUNED_TILES[n] / TUNED_KNOBS[n] for powspec_batched(tile='auto').
    """
    device = _device.require_gpu(device)
    jobs = [_Job(c, n, device) for c in catalogues]
    max_np = max(1, max(j.np for j in jobs))
    max_nb = max(j.nb for j in jobs)
    timings = {}
    best = None
    for tile in tile_candidates(n):
        b = BatchEstimator(n, len(jobs), max_np, max_nb, tile=tile, graph=graph, device=device)
        for rounds in COUNT_ROUNDS:
            for ypct in PASS_PCTS:
                for xpct in PASS_PCTS:
                    b.tune(rounds, ypct, xpct).set(jobs)
                    ms = _time_batch(b, reps)
                    timings[(tile[1], rounds, ypct, xpct)] = ms
                    if best is None or ms < best["ms"]:
                        best = {"tile": tile, "count_rounds": rounds, "ypass_pct": ypct, "xpass_pct": xpct, "ms": ms}
        del b
    TUNED_TILES[n] = best["tile"]
    TUNED_KNOBS[n] = (best["count_rounds"], best["ypass_pct"], best["xpass_pct"])
    return best, timings


def autotune_tile(n: int, catalogues, *, reps: int = 3, device=None, graph: bool = True):
    """The row-bundle height alone (default launch knobs): (1, ty, 0) and {ty: ms}."""
    device = _device.require_gpu(device)
    jobs = [_Job(c, n, device) for c in catalogues]
    max_np = max(1, max(j.np for j in jobs))
    max_nb = max(j.nb for j in jobs)
    timings = {}
    best = None
    for tile in tile_candidates(n):
        b = BatchEstimator(n, len(jobs), max_np, max_nb, tile=tile, graph=graph, device=device)
        b.set(jobs)
        timings[tile[1]] = _time_batch(b, reps)
        if best is None or timings[tile[1]] < timings[best[1]]:
            best = tile
        del b
    TUNED_TILES[n] = best
    return best, timings


def powspec_batched(catalogues, *, tile="auto", graph: bool = True, device=None, full: bool = False,
                    max_jobs_per_batch: int | None = None):
    """P_0,2,4 of every catalogue, batched by mesh side; results in input order.

    Each catalogue is a dict with x, y, z (float32 device vectors, any stride), optional w,
    box_size, n_bins and optional mas ('CIC_REF' default, 'CIC', 'TSC'), k_edges (None:
    fundamental bins), shot_noise, window, wrap, xmin/ymin/zmin.  tile: 'auto' (autotuned
    per mesh side on its own catalogues, cached), None (default rows per tile) or
    (1, ty, 0).  Every supported mesh side has a batched path (at 768 and 1024 with 16-row
    tiles and the sort's per-key tables in HBM).
    """
    device = _device.require_gpu(device)
    cats = list(catalogues)
    groups = defaultdict(list)
    for i, c in enumerate(cats):
        groups[int(c["n_bins"])].append(i)
    out = [None] * len(cats)
    runs = []
    for n, idx in sorted(groups.items()):
        sub = [cats[i] for i in idx]
        t = tile
        knobs = DEFAULT_KNOBS
        if t == "auto":
            if n not in TUNED_TILES or n not in TUNED_KNOBS:
                autotune(n, sub, device=device, graph=graph)
            t, knobs = TUNED_TILES[n], TUNED_KNOBS[n]
        jobs = [_Job(c, n, device) for c in sub]
        cap = max_jobs_per_batch or len(jobs)
        for s0 in range(0, len(jobs), cap):
            part = jobs[s0:s0 + cap]
            b = BatchEstimator(n, len(part), max(1, max(j.np for j in part)), max(j.nb for j in part), tile=t,
                               graph=graph, device=device)
            b.tune(*knobs).set(part)
            runs.append((b, idx[s0:s0 + cap]))
    run_concurrent([b for b, _ in runs])
    for b, idx in runs:
        for k, i in enumerate(idx):
            out[i] = b.result(k, full)
    torch.cuda.synchronize(device)
    return out
