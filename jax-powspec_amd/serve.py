"""PkServer: a continuously batching P_l(k) service on one GPU.

No reference counterpart (the reference is a library of functions).  Requests are
catalogues already resident in HBM (x, y, z[, w] device tensors) with their box, mesh
side, scheme, bins and shot noise.  The server keeps FFT plans and per-mesh-side engines
resident on the device:

* batch engines (mesh sides with the sorted fused path, n <= 512): a BatchEstimator whose
  slots are sized for ``max_particles``; one dispatch runs up to ``max_batch`` queued
  requests of that n with one launch per stage (HIP-graph replay once a shape repeats);
* plain engines (larger n): one pk_estimate workspace each.

Engines are created on demand while the HBM budget allows and evicted least-recently used
(idle ones only) when another mesh side needs the room.  Each engine works on its own
stream, so dispatches of different engines overlap on the device; requests that arrive
while others run are admitted as soon as an engine of their n is free (continuous
batching, not a queue drained in fixed rounds).  Latency is measured from submit() to the
result being on the host; ``stats()`` reports p50 / p99 and throughput.

    srv = PkServer(max_batch=8)
    fut = srv.submit(dict(x=x, y=y, z=z, box_size=L, n_bins=256, k_edges=edges, mas="CIC"))
    k, pk, nmodes = fut.result()      # host numpy arrays
    srv.close()
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from collections import OrderedDict, deque

import numpy as np
import torch

from . import _device, _lib, binning
from .batch import BatchEstimator, _Job
from .correlations import _Outputs

__all__ = ["PkServer", "PkFuture"]

BATCH_MAX_N = 512   # mesh sides with a batched (sorted + fused) path for both CIC and TSC


class PkFuture:
    """Result handle of one request: result() blocks until the P_l(k) is on the host."""

    def __init__(self, req: dict, t_submit: float):
        self.req = req
        self.t_submit = t_submit
        self.t_done = None
        self._ev = threading.Event()
        self._res = None
        self._err = None

    def _set(self, res, t):
        self._res = res
        self.t_done = t
        self._ev.set()

    def _fail(self, err):
        self._err = err
        self.t_done = time.perf_counter()
        self._ev.set()

    def done(self) -> bool:
        return self._ev.is_set()

    def result(self, timeout=None):
        if not self._ev.wait(timeout):
            raise TimeoutError("request not finished")
        if self._err is not None:
            raise self._err
        return self._res

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit


def _request_nb(req) -> int:
    n = int(req["n_bins"])
    if req.get("k_edges") is None:
        return binning.fundamental_grid_edges(n).size - 1
    return int(np.asarray(req["k_edges"]).size) - 1


class _BatchEngine:
    def __init__(self, n, max_jobs, max_np, max_nb, device, graph):
        self.n = n
        self.kind = "batch"
        # capacity-sized sort grids: few distinct launch shapes, so graph replays repeat
        self.est = BatchEstimator(n, max_jobs, max_np, max_nb, graph=graph, device=device, capacity_grids=True)
        self.max_jobs, self.max_np, self.max_nb = max_jobs, max_np, max_nb
        self.bytes = self.est.ws_bytes + sum(t.numel() * t.element_size() for t in
                                             (self.est.sums, self.est.counts, self.est.pk, self.est.nmodes,
                                              self.est.kmean, self.est.stats))
        self.stream = self.est._stream
        self.busy = None   # (futures, event)

    def fits(self, req):
        return int(req["x"].numel()) <= self.max_np and _request_nb(req) <= self.max_nb

    def launch(self, futs):
        st = self.stream
        for f in futs:
            st.wait_event(f.ready)
        self.est.set([f.req for f in futs], stream=st)
        self.est.run(stream=st)
        ev = torch.cuda.Event()
        ev.record(st)
        self.busy = (futs, ev)

    def collect(self):
        futs, _ = self.busy
        nb_all = [j.nb for j in self.est.jobs]
        pk = self.est.pk.cpu().numpy()
        nm = self.est.nmodes.cpu().numpy()
        km = self.est.kmean.cpu().numpy()
        out = []
        for i, f in enumerate(futs):
            j = self.est.jobs[i]
            nb = nb_all[i]
            kF = 2.0 * np.pi / j.box
            k = km[i, :nb] * kF if j.fundamental else 0.5 * (j.edges[1:] + j.edges[:-1]) * kF
            out.append((f, (k.astype(np.float32), pk[i, :nb].copy(), nm[i, :nb].copy())))
        self.busy = None
        return out


class _PlainEngine:
    def __init__(self, n, max_np, max_nb, device):
        self.n = n
        self.kind = "plain"
        self.device = device
        self.max_np, self.max_nb = max_np, max_nb
        self.plan = _device.plan(n, device)
        lib = _lib.load()
        nbytes = C.c_size_t()
        with torch.cuda.device(device):
            # TSC needs the most scratch; size for it so any scheme fits
            _lib.check("pk_estimate_workspace_size", lib.pk_estimate_workspace_size(
                self.plan.handle, max_np, _lib.STENCIL["TSC"], None, max_nb, C.byref(nbytes)))
            need = nbytes.value
            for st in ("CIC_REF", "CIC"):
                _lib.check("pk_estimate_workspace_size", lib.pk_estimate_workspace_size(
                    self.plan.handle, max_np, _lib.STENCIL[st], None, max_nb, C.byref(nbytes)))
                need = max(need, nbytes.value)
        self.ws_bytes = need
        self.ws = _device.workspace(need, device)
        self.out = _Outputs(max_nb, device, True)
        self.stats_t = torch.zeros(_lib.STATS_LEN, dtype=torch.float64, device=device)
        self.bytes = need + max_nb * (24 + 8 + 12 + 4 + 8)
        self.stream = torch.cuda.Stream(device=device)
        self.busy = None

    def fits(self, req):
        return int(req["x"].numel()) <= self.max_np and _request_nb(req) <= self.max_nb

    def launch(self, futs):
        (f,) = futs
        st = self.stream
        st.wait_event(f.ready)
        j = _Job(f.req, self.n, self.device)
        o = self.out
        with torch.cuda.device(self.device):
            _lib.check("pk_estimate", _lib.load().pk_estimate(
                self.plan.handle, _device.ptr(j.xs), j.sx, _device.ptr(j.ys), j.sy, _device.ptr(j.zs), j.sz,
                _device.ptr(j.ws), j.sw, j.np, *j.mins, j.box, _lib.STENCIL[j.mas], j.wrap, None,
                _device.ptr(j.tables.thr), _device.ptr(j.tables.jst), _device.ptr(j.tables.modes),
                _device.ptr(j.tables.ksum), j.nb, j.p, j.flags, j.shot, _device.ptr(o.sums),
                _device.ptr(o.counts), _device.ptr(o.pk), _device.ptr(o.nmodes), _device.ptr(o.kmean),
                _device.ptr(self.stats_t), _device.ptr(self.ws), self.ws_bytes, st.cuda_stream))
        ev = torch.cuda.Event()
        ev.record(st)
        self.busy = (futs, ev, j)

    def collect(self):
        futs, _, j = self.busy
        nb = j.nb
        kF = 2.0 * np.pi / j.box
        pk = self.out.pk[:nb].cpu().numpy().copy()
        nm = self.out.nmodes[:nb].cpu().numpy().copy()
        if j.fundamental:
            k = self.out.kmean[:nb].cpu().numpy() * kF
        else:
            k = 0.5 * (j.edges[1:] + j.edges[:-1]) * kF
        self.busy = None
        return [(futs[0], (k.astype(np.float32), pk, nm))]


class PkServer:
    """Continuously batching P_l(k) server (see the module docstring)."""

    def __init__(self, *, device=None, hbm_budget: int | None = None, max_batch: int = 8, engines_per_n: int = 2,
                 max_particles: int = 30_000_000, graph: bool = True):
        self.device = _device.require_gpu(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.6 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.engines_per_n = int(engines_per_n)
        self.max_particles = int(max_particles)
        self.graph = graph
        self.reserved = 0
        self.engines: "OrderedDict[int, _BatchEngine | _PlainEngine]" = OrderedDict()   # LRU order
        self._next_id = 0
        self.queues: "dict[int, deque]" = {}   # per mesh side, arrival order
        self.lock = threading.Lock()
        self.wake = threading.Condition(self.lock)
        self.completed: list = []
        self.errors = 0
        self.closed = False
        self.thread = threading.Thread(target=self._loop, name="PkServer", daemon=True)
        self.thread.start()

    # -- public --------------------------------------------------------------------------
    def submit(self, req: dict) -> PkFuture:
        n = int(req["n_bins"])
        if not _lib.load().pk_supported_n(n):
            raise ValueError(f"unsupported mesh side {n}")
        if int(req["x"].numel()) > self.max_particles:
            raise ValueError(f"{req['x'].numel()} particles exceed the server's max_particles={self.max_particles}")
        if _request_nb(req) > self._engine_size(n)[1]:
            raise ValueError(f"{_request_nb(req)} bins exceed the server's capacity for n={n}")
        f = PkFuture(req, time.perf_counter())
        f.ready = torch.cuda.Event()
        f.ready.record(torch.cuda.current_stream(self.device))   # the catalogue is complete here
        with self.wake:
            if self.closed:
                raise RuntimeError("server is closed")
            self.queues.setdefault(n, deque()).append(f)
            self.wake.notify()
        return f

    def close(self, wait: bool = True):
        with self.wake:
            self.closed = True
            self.wake.notify()
        if wait:
            self.thread.join()

    def stats(self):
        lat = np.array([f.latency for f in self.completed if f.latency is not None])
        if lat.size == 0:
            return {}
        t0 = min(f.t_submit for f in self.completed)
        t1 = max(f.t_done for f in self.completed)
        return {"requests": int(lat.size), "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
                "throughput_per_s": float(lat.size / max(t1 - t0, 1e-9)), "errors": self.errors,
                "engines": {str(k): (e.kind, e.n) for k, e in self.engines.items()},
                "reserved_bytes": self.reserved}

    # -- scheduler -----------------------------------------------------------------------
    def _engine_size(self, n):
        # capacity of a new engine: the server's particle cap; bins up to the fundamental
        # binning of n or n / 2 + 1 user bins, whichever is more
        nb = max(binning.fundamental_grid_edges(n).size - 1, n // 2 + 1)
        return self.max_particles, nb

    def _make_engine(self, n):
        max_np, max_nb = self._engine_size(n)
        with torch.cuda.device(self.device):
            if n <= BATCH_MAX_N:
                return _BatchEngine(n, self.max_batch, max_np, max_nb, self.device, self.graph)
            return _PlainEngine(n, max_np, max_nb, self.device)

    def _estimate_bytes(self, n):
        max_np, max_nb = self._engine_size(n)
        plan = _device.plan(n, self.device)
        lib = _lib.load()
        nbytes = C.c_size_t()
        if n <= BATCH_MAX_N:
            _lib.check("pk_batch_workspace_size", lib.pk_batch_workspace_size(
                plan.handle, self.max_batch, max_np, max_nb, None, C.byref(nbytes)))
        else:
            _lib.check("pk_estimate_workspace_size", lib.pk_estimate_workspace_size(
                plan.handle, max_np, _lib.STENCIL["TSC"], None, max_nb, C.byref(nbytes)))
        return nbytes.value

    def _free_engine_for(self, n):
        """An idle engine of mesh side n, creating one (evicting idle LRU engines of other
        sides) while the budget allows; None if every engine of n is busy or memory is short."""
        mine = [(k, e) for k, e in self.engines.items() if e.n == n]
        for k, e in mine:
            if e.busy is None:
                self.engines.move_to_end(k)
                return e
        if len(mine) >= self.engines_per_n:
            return None
        need = self._estimate_bytes(n)
        if need > self.hbm_budget:
            raise MemoryError(f"one engine for n={n} needs {need} bytes > budget {self.hbm_budget}")
        while self.reserved + need > self.hbm_budget:
            victim = next((k for k, e in self.engines.items() if e.busy is None and e.n != n), None)
            if victim is None:
                return None   # memory frees up when a running dispatch completes
            e = self.engines.pop(victim)
            self.reserved -= e.bytes
            del e
        e = self._make_engine(n)
        key = self._next_id
        self._next_id += 1
        self.engines[key] = e
        self.reserved += e.bytes
        return e

    def _dispatch(self):
        """Admit queued requests onto free engines, the mesh side whose oldest request has
        waited longest first; a batch engine takes up to max_batch requests of its n."""
        order = sorted((q[0].t_submit, n) for n, q in self.queues.items() if q)
        for _, n in order:
            q = self.queues[n]
            while q:
                try:
                    e = self._free_engine_for(n)
                except Exception as err:   # can never run (e.g. larger than the budget)
                    while q:
                        f = q.popleft()
                        self.errors += 1
                        f._fail(err)
                    break
                if e is None:
                    break
                take = []
                while q and len(take) < (e.max_jobs if e.kind == "batch" else 1):
                    take.append(q.popleft())
                try:
                    e.launch(take)
                except Exception as err:
                    e.busy = None
                    for g in take:
                        self.errors += 1
                        g._fail(err)

    def _poll(self):
        done_any = False
        for e in list(self.engines.values()):
            if e.busy is not None and e.busy[1].query():
                for f, res in e.collect():
                    f._set(res, time.perf_counter())
                    self.completed.append(f)
                done_any = True
        return done_any

    def _loop(self):
        torch.cuda.set_device(self.device)
        while True:
            with self.wake:
                busy = any(e.busy is not None for e in self.engines.values())
                if not any(self.queues.values()) and not busy:
                    if self.closed:
                        return
                    self.wake.wait(timeout=0.05)
                    continue
                self._poll()
                self._dispatch()
                busy = any(e.busy is not None for e in self.engines.values())
            if busy:
                time.sleep(0.0002)
