"""Benchmark: P_0,2,4(k) estimates/s at 512^3 with ~2e7 particles on one MI355X.

Workload (synthetic, generated on the GPU with torch before timing; seeds fixed):
  lognormal mock in L = 1500 Mpc/h: 512^3 Gaussian field with
  P(k) = 2e4 (k/0.05)^-1.5 exp(-(k/0.5)^2), lognormal transform, Poisson sampling at
  nbar = 2e7 / L^3 (N_p ~ 2.0e7), uniform positions within cells, redshift-space
  displacement z += N(0, 4 Mpc/h) wrapped, catalogue shuffled (arbitrary order).
  One step = one full estimate: reference-stencil CIC assignment -> contrast -> R2C FFT
  -> CIC deconvolution -> P_0,2,4 in 256 k-bins with edges k_F (0.5 + i), i = 0..256.

Prints one JSON line with throughput, the roofline of the dominant kernel (from HIP events
recorded around each stage on the launch stream during the timed steps) and a CPU
baseline (FP32 NumPy/SciPy restatement of the reference algorithm on this host).
"""
from __future__ import annotations

import argparse
import math
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_MESH = 512
BOX = 1500.0
NBAR = 2.0e7 / BOX ** 3
NBINS = 256


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_catalogue(n, box, nbar, device, seed=2002):
    """Lognormal + Poisson + RSD mock on the GPU (input generation only)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    kf = 2 * np.pi / box
    ki = torch.fft.fftfreq(n, d=1.0 / n, device=device)
    kz = torch.fft.rfftfreq(n, d=1.0 / n, device=device)
    k = kf * torch.sqrt(ki[:, None, None] ** 2 + ki[None, :, None] ** 2 + kz[None, None, :] ** 2)
    pk = 2e4 * (k.clamp_min(1e-12) / 0.05) ** -1.5 * torch.exp(-(k / 0.5) ** 2)
    pk[0, 0, 0] = 0.0
    vcell = (box / n) ** 3
    white = torch.randn((n, n, n), device=device, generator=g)
    wk = torch.fft.rfftn(white)
    wk *= torch.sqrt(pk / vcell)
    del pk, k
    dg = torch.fft.irfftn(wk, s=(n, n, n))
    del wk, white
    sig2 = dg.var()
    rho = torch.exp(dg - sig2 / 2)
    rho /= rho.mean()
    del dg
    lam = (rho * (nbar * vcell)).reshape(-1)
    counts = torch.poisson(lam, generator=g).to(torch.int64)
    del lam, rho
    cells = torch.repeat_interleave(torch.arange(n ** 3, device=device), counts)
    npart = cells.numel()
    h = box / n
    ix = cells // (n * n)
    iy = (cells // n) % n
    iz = cells % n
    pos = torch.empty((npart, 3), device=device, dtype=torch.float32)
    u = torch.rand((npart, 3), device=device, generator=g)
    pos[:, 0] = (ix + u[:, 0]) * h
    pos[:, 1] = (iy + u[:, 1]) * h
    pos[:, 2] = torch.remainder((iz + u[:, 2]) * h + 4.0 * torch.randn(npart, device=device, generator=g), box)
    del u, cells, ix, iy, iz
    perm = torch.randperm(npart, device=device, generator=g)
    pos = pos[perm].contiguous()
    pos.clamp_(0.0, np.nextafter(np.float32(box), np.float32(0)))
    return pos


def algorithmic_bytes(n, npart, fused, rec=12):
    """Minimum HBM bytes each stage must move per launch (rec = bytes per sorted record:
    12 for unit weights, 16 with weights)."""
    half = 8 * n * n * (n // 2 + 1)
    ab = {
        "sort_count": 12 * npart,                  # read positions
        "sort_place": 12 * npart + rec * npart,    # read positions, write plane-grouped records
        "sort_fine": 2 * rec * npart,              # read them, write the key-sorted records
        "ypass": 2 * half,                         # in place
        "xpass_bin": half,                         # read; only nb-sized sums leave
    }
    if fused:   # deposit fused with the z pass: read records, write the half spectrum
        ab["deposit"] = rec * npart + half
    else:
        ab["deposit"] = rec * npart + 4 * n ** 3
        ab["zpass"] = 4 * n ** 3 + half
    return ab


# stage -> kernel-name fragment in rocprofv3 output
STAGE_KERNELS = {
    "sort_count": "k_bucket_count", "sort_place": "k_bucket_place", "sort_fine": "k_fine_place", "ypass": "k_ypass<",
    "xpass_bin": "k_xpass_bin<", "zpass": "k_zpass<",
}


def stage_kernel(stage, n, fused, stencil=0):
    if stage == "deposit":
        # default row-bundle height of the fused kernel (largest of 32, 24, 16 dividing n,
        # 16 at 768 and 1024) and the stencil it is instantiated for (0: reference CIC)
        ty = n if n <= 32 else (32 if n % 32 == 0 else (24 if n % 24 == 0 else 16))
        if ty * n * 8 > 131072:
            ty = 16
        return f"k_fused_deposit_zpass<{n}, {ty}, {stencil}>" if fused else "k_rows_deposit"
    return STAGE_KERNELS.get(stage, stage)


def trace_summary(stats_csv, out_path, title, command):
    """The library's kernels from a rocprofv3 --kernel-trace --stats kernel_stats.csv."""
    import csv
    rows = [r for r in csv.DictReader(open(stats_csv)) if r["Name"].startswith(("pk::", "void pk::"))]
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = [f"# rocprofv3 kernel-trace statistics: {title}", f"# command: {command}",
             "# kernel\tcalls\tavg_us\tmin_us\tmax_us\ttotal_%"]
    for r in rows:
        name = r["Name"].split("(")[0].replace("void ", "")
        lines.append(f"{name}\t{r['Calls']}\t{float(r['AverageNs']) / 1e3:.1f}\t{float(r['MinNs']) / 1e3:.1f}\t"
                     f"{float(r['MaxNs']) / 1e3:.1f}\t{float(r['Percentage']):.2f}")
    open(out_path, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def read_traffic(path):
    """HBM bytes per dispatch from a traffic summary written by --pmc-summary:
    lines 'kernel<TAB>fetch_bytes<TAB>write_bytes' after the header lines starting with '#'."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        k, f, w = line.rstrip("\n").split("\t")[:3]
        out[k] = (float(f), float(w))
    return out


def pmc_summary(fetch_csv, write_csv, out_path, command):
    """Per-kernel mean FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from two rocprofv3
    --pmc runs, calibrated with the 1 GiB device copy of --calib-copy (the first copy
    dispatches: 1 GiB read, 1 GiB written)."""
    import collections
    import csv

    def load(path, ctr):
        agg = collections.defaultdict(list)
        copies = []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != ctr:
                continue
            name = r["Kernel_Name"]
            v = float(r["Counter_Value"]) * 1024.0
            if "copyBuffer" in name and int(r["Grid_Size"]) == 131072 and len(copies) < 3:
                copies.append(v)
            agg[name].append(v)
        return agg, copies

    fa, fc = load(fetch_csv, "FETCH_SIZE")
    wa, wc = load(write_csv, "WRITE_SIZE")
    gib = float(1 << 30)
    ff = gib / (sum(fc) / len(fc))
    fw = gib / (sum(wc) / len(wc))
    lines = [f"# HBM traffic per dispatch (bytes), rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE",
             f"# raw counters: {command}",
             f"# calibration: 1 GiB device copy reads FETCH_SIZE x {ff:.3f}, writes WRITE_SIZE x {fw:.3f}",
             "# kernel\tfetch_bytes\twrite_bytes\tdispatches"]
    for name in sorted(fa):
        if not name.startswith("void pk::") and not name.startswith("pk::"):
            continue
        f = sum(fa[name]) / len(fa[name]) * ff
        w = sum(wa.get(name, [0.0])) / max(1, len(wa.get(name, []))) * fw
        short = name.split("(")[0].replace("void ", "").replace("pk::", "")
        lines.append(f"{short}\t{f:.0f}\t{w:.0f}\t{len(fa[name])}")
    open(out_path, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def cpu_baseline(pos_host, n, box, edges, shot, reps=5):
    """float32 CPU restatement of the reference (oracle/cpu_fp32.py) on the same catalogue:
    one warm-up, then the median of `reps` timed estimates."""
    from oracle import cpu_fp32
    cores = cpu_fp32.ncores()
    run = lambda: cpu_fp32.estimate_f32(pos_host[:, 0], pos_host[:, 1], pos_host[:, 2], box, n, edges, shot,
                                        workers=cores)
    run()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    return {"value": 1.0 / dt, "unit": "estimates/s", "cores": cores, "kind": "port",
            "sample": f"median of {reps} full estimates (after 1 warm-up) of the same {n}^3 / "
                      f"{pos_host.shape[0]:.3g}-particle catalogue ({dt:.1f} s each, range {min(ts):.1f}-"
                      f"{max(ts):.1f} s; FFT on {cores} threads, scatter and binning single-threaded NumPy)"}


def parity_vs_oracle(pos_host, n, box, edges, shot, sums, counts):
    """Per-multipole error of the timed GPU estimate against the FP64 oracle on the same
    float32 positions (checker only): eps_l = max_b |P_l - P_l^oracle| / A_l,b with A the
    same-bin scale sum |L_l| P2 norm (2l+1) / N_b, and the median plain relative error."""
    from oracle import ref_fp64 as R
    t0 = time.perf_counter()
    res = R.estimate(pos_host[:, 0], pos_host[:, 1], pos_host[:, 2], None, box, n, edges, stencil="CIC_REF", p=2)
    opk, oex = np.asarray(res[1], np.float64), res[3]
    cnt = np.asarray(counts)
    P = np.asarray(sums) / cnt[:, None] * (box / n ** 2) ** 3 * np.array([1.0, 5.0, 9.0])
    eps = np.abs(P - opk) / oex["scale"]
    rel = np.abs(P - opk) / np.abs(opk)
    return {"oracle": "FP64 NumPy restatement of the reference (oracle/ref_fp64.py), same float32 positions",
            "eps_same_bin_scale": {f"P{2 * l}": float(eps[:, l].max()) for l in range(3)},
            "rel_median": {f"P{2 * l}": float(np.median(rel[:, l])) for l in range(3)},
            "rel_max": {f"P{2 * l}": float(rel[:, l].max()) for l in range(3)},
            "gate": 1e-5, "counts_equal": bool(np.array_equal(cnt, oex["counts"])),
            "oracle_s": round(time.perf_counter() - t0, 1)}


def bench_small(dev, steps=50, cpu=True):
    """Plain small case: 1e6 uniform particles (seed 1001) in L = 1000 Mpc/h, 128^3 CIC
    (reference stencil), monopole in k_F bins from k_F to k_N, shot noise 1/nbar."""
    import torch
    from jax_powspec_amd.correlations import Estimator
    n, box, npart = 128, 1000.0, 1_000_000
    pos_h = (np.random.default_rng(1001).random((npart, 3), dtype=np.float32) * np.float32(box)).astype(np.float32)
    pos = torch.as_tensor(pos_h, device=dev)
    kf = 2 * np.pi / box
    edges = kf * np.arange(1, n // 2 + 1)
    shot = box ** 3 / npart
    est = Estimator(box_size=box, n_bins=n, max_particles=npart, k_edges=edges, mas="CIC_REF", shot_noise=shot,
                    device=dev)
    for _ in range(3):
        est.run(pos[:, 0], pos[:, 1], pos[:, 2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        est.run(pos[:, 0], pos[:, 1], pos[:, 2])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    out = {"workload": "1e6 uniform particles, L=1000 Mpc/h, 128^3 CIC (reference stencil), P_0 in 63 k_F bins",
           "estimates_per_s": round(1e3 / ms, 1), "ms_per_estimate": round(ms, 4)}
    if cpu:
        from oracle import cpu_fp32
        cores = cpu_fp32.ncores()
        cpu_fp32.estimate_f32(pos_h[:, 0], pos_h[:, 1], pos_h[:, 2], box, n, edges, shot, workers=cores)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            cpu_fp32.estimate_f32(pos_h[:, 0], pos_h[:, 1], pos_h[:, 2], box, n, edges, shot, workers=cores)
        dt = (time.perf_counter() - t0) / reps
        out["cpu_estimates_per_s"] = round(1.0 / dt, 3)
        out["cpu_cores"] = cores
        out["gpu_over_cpu"] = round(dt * 1e3 / ms, 1)
    return out


BATCH_MOCKS = 256
BATCH_NS = (64, 96, 128, 192, 256)


def make_batch_mocks(dev, count=BATCH_MOCKS, seed=3000):
    """Covariance set of uneven lognormal mocks: L ~ U[500, 1500] Mpc/h, n ~ {64..256},
    N_p log-uniform in [1e5, 4e6], CIC for even i and TSC for odd i, k_F bins up to k_N."""
    cats = []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        box = float(rng.uniform(500.0, 1500.0))
        n = int(rng.choice(BATCH_NS))
        npart = float(np.exp(rng.uniform(np.log(1e5), np.log(4e6))))
        pos = make_catalogue(n, box, npart / box ** 3, dev, seed=seed + i)
        kf = 2 * np.pi / box
        cats.append(dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], box_size=box, n_bins=n,
                         mas="CIC" if i % 2 == 0 else "TSC", k_edges=kf * np.arange(1, n // 2 + 1),
                         shot_noise=box ** 3 / pos.shape[0]))
    return cats


def bench_batched(dev, reps=5):
    """256 uneven mocks: plain one-call-per-catalogue estimate() vs resident graph-replayed
    batches with autotuned tiles.  Both produce the same P_l (checked here)."""
    import torch
    from jax_powspec_amd import batch, correlations
    log("batched: generating mocks ...")
    cats = make_batch_mocks(dev)
    npart = sum(int(c["x"].numel()) for c in cats)
    torch.cuda.synchronize()

    def plain_pass():
        return [correlations.estimate(c["x"], c["y"], c["z"], box_size=c["box_size"], n_bins=c["n_bins"],
                                      k_edges=c["k_edges"], mas=c["mas"], shot_noise=c["shot_noise"], full=True)
                for c in cats]

    plain_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ref = plain_pass()
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / reps
    log(f"batched: plain {plain_s * 1e3:.1f} ms for {len(cats)} mocks")
    groups = {}
    for i, c in enumerate(cats):
        groups.setdefault(c["n_bins"], []).append(i)
    chosen, tune = {}, {}
    batches = []
    for n, idx in sorted(groups.items()):
        sub = [cats[i] for i in idx]
        best, timings = batch.autotune(n, sub, reps=2, device=dev)
        chosen[n] = {"ty": best["tile"][1], "count_rounds": best["count_rounds"], "ypass_pct": best["ypass_pct"],
                     "xpass_pct": best["xpass_pct"], "ms": round(best["ms"], 3),
                     "default_ms": round(timings[(best["tile"][1],) + batch.DEFAULT_KNOBS], 3)}
        tune[n] = len(timings)
        b = batch.BatchEstimator(n, len(sub), max(int(c["x"].numel()) for c in sub),
                                 max(len(c["k_edges"]) - 1 for c in sub), tile=best["tile"], graph=True, device=dev)
        b.tune(best["count_rounds"], best["ypass_pct"], best["xpass_pct"]).set(sub)
        batches.append((b, idx))
    bl = [b for b, _ in batches]
    for _ in range(2):   # plain launches, then capture
        batch.run_concurrent(bl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for b in bl:
            b.run()
    torch.cuda.synchronize()
    batch_serial_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.run_concurrent(bl)
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        for b, idx in batches:
            b.set([cats[i] for i in idx])
        batch.run_concurrent(bl)
    torch.cuda.synchronize()
    batch_set_s = (time.perf_counter() - t0) / reps
    # same results as the plain path (same tiles would be bit-identical up to FP64 sum order)
    worst = 0.0
    counts_equal = True
    for b, idx in batches:
        for k, i in enumerate(idx):
            _, pk_b, nm_b, ex_b = b.result(k, full=True)
            _, pk_p, nm_p, ex_p = ref[i]
            counts_equal &= bool(torch.equal(ex_b["counts"], ex_p["counts"]))
            s_b, s_p = ex_b["sums"], ex_p["sums"]
            scale = s_p.abs().amax(dim=0, keepdim=True).clamp_min(1e-300)
            worst = max(worst, float(((s_b - s_p).abs() / scale).max()))
    log(f"batched: {batch_s * 1e3:.1f} ms per pass ({plain_s / batch_s:.2f}x), tuned {chosen}")
    return {"mocks": len(cats), "particles": npart, "mesh_sides": {str(n): len(v) for n, v in sorted(groups.items())},
            "plain_ms": round(plain_s * 1e3, 3), "batched_ms": round(batch_s * 1e3, 3),
            "batched_with_refill_ms": round(batch_set_s * 1e3, 3),
            "batched_groups_in_sequence_ms": round(batch_serial_s * 1e3, 3),
            "plain_estimates_per_s": round(len(cats) / plain_s, 1),
            "batched_estimates_per_s": round(len(cats) / batch_s, 1),
            "speedup": round(plain_s / batch_s, 3), "autotuned": {str(k): v for k, v in chosen.items()},
            "autotune_configs_timed": {str(k): v for k, v in tune.items()},
            "autotune_space": {"ty": "compiled fused variants", "count_rounds": list(batch.COUNT_ROUNDS),
                               "ypass_pct": list(batch.PASS_PCTS), "xpass_pct": list(batch.PASS_PCTS)},
            "max_rel_sum_diff_vs_plain": worst, "counts_equal": counts_equal,
            "note": "plain = correlations.estimate() per mock; batched = one graph-replayed launch sequence "
                    "per mesh side over resident slots, the mesh-side groups on concurrent streams; "
                    "inputs resident in HBM"}


SERVE_REQUESTS = 2000
SERVE_NS = (128, 256, 384, 512, 768, 1024)


def make_serve_requests(dev, count=SERVE_REQUESTS, seed=5000):
    """Requests of the serving benchmark: each draws n from {128, ..., 1024} and N_p
    log-uniform in [1e5, 3e7]; its particles are the first N_p of one shuffled lognormal
    catalogue (L = 1500 Mpc/h, 3e7 particles), i.e. a random subsample of the same field."""
    pool = make_catalogue(512, BOX, 3.0e7 / BOX ** 3, dev, seed=seed)
    reqs = []
    kf = 2 * np.pi / BOX
    for i in range(count):
        rng = np.random.default_rng(seed + 1 + i)
        n = int(rng.choice(SERVE_NS))
        npart = min(int(np.exp(rng.uniform(np.log(1e5), np.log(3e7)))), pool.shape[0])
        p = pool[:npart]
        reqs.append(dict(x=p[:, 0], y=p[:, 1], z=p[:, 2], box_size=BOX, n_bins=n, mas="CIC_REF",
                         k_edges=kf * (0.5 + np.arange(n // 2 + 1)), shot_noise=BOX ** 3 / npart))
    return pool, reqs


def bench_served(dev, plain_sample=300):
    """Serving benchmark: sequential plain estimate() calls vs PkServer, at saturation and
    under Poisson arrivals at 0.8x the plain path's measured capacity."""
    import torch
    from jax_powspec_amd import correlations
    from jax_powspec_amd.serve import PkServer
    log("served: generating requests ...")
    pool, reqs = make_serve_requests(dev)
    torch.cuda.synchronize()

    def plain(r):
        k, pk, nm = correlations.estimate(r["x"], r["y"], r["z"], box_size=r["box_size"], n_bins=r["n_bins"],
                                          k_edges=r["k_edges"], mas=r["mas"], shot_noise=r["shot_noise"])
        return pk.cpu().numpy()

    for n in SERVE_NS:   # warm plans and kernels
        plain(next(r for r in reqs if r["n_bins"] == n))
    torch.cuda.synchronize()
    sample = reqs[:plain_sample]
    t0 = time.perf_counter()
    lat_plain = []
    for r in sample:
        t = time.perf_counter()
        plain(r)
        lat_plain.append(time.perf_counter() - t)
    plain_s = time.perf_counter() - t0
    plain_rate = len(sample) / plain_s
    log(f"served: plain {plain_rate:.1f} req/s on {len(sample)} requests")
    srv = PkServer(device=dev, max_batch=8)
    try:
        for n in SERVE_NS:   # warm the engines (and their graphs) once
            srv.submit(next(r for r in reqs if r["n_bins"] == n)).result(timeout=300)
        srv.completed.clear()
        futs = [srv.submit(r) for r in reqs]
        for f in futs:
            f.result(timeout=600)
        sat = srv.stats()
        log(f"served: saturation {sat['throughput_per_s']:.1f} req/s")
        srv.completed.clear()
        lam = 0.8 * plain_rate
        rng = np.random.default_rng(6000)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(reqs)))
        t0 = time.perf_counter()
        futs = []
        for r, ta in zip(reqs, arrivals):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            futs.append(srv.submit(r))
        for f in futs:
            f.result(timeout=600)
        poi = srv.stats()
        log(f"served: poisson {poi['throughput_per_s']:.1f} req/s, p50 {poi['p50_ms']:.2f} ms, p99 {poi['p99_ms']:.2f} ms")
        # the same requests' results alone (plain) and from the server agree
        worst = 0.0
        for f in futs[:40]:
            want = plain(f.req)
            got = f.result()[1]
            ok = ~np.isnan(want)
            scale = np.nanmax(np.abs(want), axis=0, keepdims=True)   # per multipole
            worst = max(worst, float(np.max((np.abs(got - want) / scale)[ok])))
    finally:
        srv.close()
    lp = np.array(lat_plain) * 1e3
    return {"requests": len(reqs), "mesh_sides": list(SERVE_NS), "particles_range": [1e5, 3e7],
            "plain_sequential_req_per_s": round(plain_rate, 2),
            "plain_p50_ms": round(float(np.percentile(lp, 50)), 3), "plain_p99_ms": round(float(np.percentile(lp, 99)), 3),
            "plain_sample": len(sample),
            "served_saturation_req_per_s": round(sat["throughput_per_s"], 2),
            "served_saturation_speedup": round(sat["throughput_per_s"] / plain_rate, 3),
            "served_saturation_p50_ms": round(sat["p50_ms"], 3), "served_saturation_p99_ms": round(sat["p99_ms"], 3),
            "poisson_rate_req_per_s": round(lam, 2), "poisson_throughput_req_per_s": round(poi["throughput_per_s"], 2),
            "poisson_p50_ms": round(poi["p50_ms"], 3), "poisson_p99_ms": round(poi["p99_ms"], 3),
            "max_diff_vs_plain_first40_rel_to_multipole_max": worst, "errors": sat["errors"] + poi["errors"],
            "note": "latency = submit to P_l on the host; catalogues resident in HBM; plain = correlations.estimate() "
                    "one request at a time (capacity measured on the first requests)"}


MOCK_KF = np.linspace(1e-4, 2.0, 4056)
MOCK_PKF = 2e4 * (MOCK_KF / 0.05) ** -1.5 * np.exp(-(MOCK_KF / 0.5) ** 2)


def mock_bytes(n, npart, rsd):
    """Minimum HBM bytes of one mock: field pass (write half spectrum), inverse y (read +
    write), C2R (read half spectrum, write mesh), three population reads of the field,
    one-byte counts written and read, cell ids written and read, records written; with
    RSD a second field pass and the psi_z reads."""
    half = 8 * n * n * (n // 2 + 1)
    mesh = 4 * n ** 3
    field = half + 2 * half + half + mesh
    pop = 3 * mesh + 2 * n ** 3 + 8 * npart + 12 * npart
    return (2 if rsd else 1) * field + pop + (4 * npart if rsd else 0)


def bench_mock(dev, reps=5, cpu=True):
    """Seeded lognormal mock at the headline size (tests/create_lognormal.py:42-55 on the
    GPU): Gaussian field of P(k) -> irfftn -> exp(b g) -> Poisson population, 512^3 cells,
    L = 1500 Mpc/h, 2e7 particles on average; plus the Zel'dovich redshift-space shift."""
    import torch
    from jax_powspec_amd import gauss_field as GF

    n, box, bias, f = N_MESH, BOX, 1.1, 0.7
    out = {"workload": f"lognormal mock {n}^3 cells, L={BOX:g} Mpc/h, nbar={NBAR:.3e}, b={bias}, "
                       f"f={f} (RSD run)", "mesh": n}
    for growth in (0.0, f):
        GF.lognormal_mock(n, MOCK_KF, MOCK_PKF, box, NBAR, bias, 11, growth_rate=growth)
        torch.cuda.synchronize()
        ts = []
        for r in range(reps):
            t0 = time.perf_counter()
            pos = GF.lognormal_mock(n, MOCK_KF, MOCK_PKF, box, NBAR, bias, 100 + r, growth_rate=growth)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts)) * 1e3
        npart = int(pos.shape[0])
        key = "rsd" if growth else "real"
        gbs = mock_bytes(n, npart, bool(growth)) / (ms * 1e-3) / 1e9
        out[key] = {"ms": round(ms, 3), "particles": npart, "particles_per_s": round(npart / (ms * 1e-3), 1),
                    "algorithmic_gbs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
        del pos
        torch.cuda.empty_cache()
    if cpu:
        # the oracle's NumPy pipeline on a 128^3 sample of the same cell size, P(k), density
        # and bias (same rates per cell), per-cell time scaled to 512^3
        from oracle import mocks as M
        ns = 128
        bs = box * ns / n
        t0 = time.perf_counter()
        d = M.gaussian_field(ns, MOCK_KF, MOCK_PKF, 0, 11, bs)
        g = M.irfftn(d).astype(np.float32)
        p, _ = M.populate(g, ns, bs, NBAR, 12, bias=bias)
        dt = time.perf_counter() - t0
        scaled = dt * (n / ns) ** 3
        out["cpu_baseline"] = {"ms_scaled": round(scaled * 1e3, 1), "unit": "ms per mock", "cores": 1,
                               "kind": "port", "extrapolated": True,
                               "sample": f"oracle NumPy pipeline at {ns}^3 (L={bs:g}, {p.shape[0]} particles, "
                                         f"{dt:.2f} s), per-cell time scaled to {n}^3"}
        out["gpu_over_cpu"] = round(scaled * 1e3 / out["real"]["ms"], 1)
    return out


def xi_bytes(n):
    """Minimum HBM bytes of one xi_vec: z pass (read mesh, write half spectrum), forward y
    (read + write), x pass with power and inverse (read + write), inverse y (read + write),
    C2R with binning (read half spectrum)."""
    half = 8 * n * n * (n // 2 + 1)
    return 4 * n ** 3 + half + 2 * half + 2 * half + 2 * half + half


def bench_xi(dev, reps=10, cpu=True):
    """xi_vec (src/correlations.py:119-187) of a 512^3 Gaussian field with the separation
    bins of tests/create_lognormal.py:66 (s_edges = arange(1e-3, 200, 5), L = 1500)."""
    import torch
    from jax_powspec_amd import correlations, gauss_field as GF

    n, box = N_MESH, BOX
    edges = np.arange(1e-3, 200.0, 5.0)
    d = GF.gaussian_mesh(n, MOCK_KF, MOCK_PKF, 1, 21, box)
    correlations.xi_vec(d, box, edges)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r, xi, nm = correlations.xi_vec(d, box, edges)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ms = float(np.median(ts)) * 1e3
    gbs = xi_bytes(n) / (ms * 1e-3) / 1e9
    out = {"workload": f"xi_vec {n}^3, L={box:g}, s_edges=arange(1e-3, 200, 5) ({edges.size - 1} bins)",
           "ms": round(ms, 3), "per_s": round(1e3 / ms, 2), "algorithmic_gbs": round(gbs, 1),
           "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
    del d
    torch.cuda.empty_cache()
    if cpu:
        from oracle import ref_fp64 as R
        ns = 128
        dh = np.random.default_rng(0).standard_normal((ns, ns, ns)).astype(np.float32)
        t0 = time.perf_counter()
        R.xi_vec(dh, box * ns / n, edges * ns / n)
        dt = time.perf_counter() - t0
        scaled = dt * (n / ns) ** 3 * math.log2(n) / math.log2(ns)
        out["cpu_baseline"] = {"ms_scaled": round(scaled * 1e3, 1), "unit": "ms per xi_vec", "cores": 1,
                               "kind": "port", "extrapolated": True,
                               "sample": f"oracle FP64 NumPy xi_vec at {ns}^3 ({dt:.2f} s), scaled by n^3 log n "
                                         f"to {n}^3"}
        out["gpu_over_cpu"] = round(scaled * 1e3 / ms, 1)
    return out


def bench_bispec(dev, reps=3, cpu=True):
    """bispec (src/correlations.py:334-458) of a 512^3 lognormal field with the triangles of
    tests/bispec.py:53-54 (k1 = 0.1, k2 = 0.2, 20 angles; L = 1500)."""
    import torch
    from jax_powspec_amd import correlations, gauss_field as GF

    n, box = N_MESH, BOX
    k1, k2 = 0.1, 0.2
    theta = np.linspace(0, np.pi, 20)
    g = GF.gaussian_mesh(n, MOCK_KF, MOCK_PKF, 1, 31, box)
    d = torch.exp(0.5 * g / g.std())
    d = d / d.mean() - 1.0
    del g
    def timed(reuse):
        correlations.bispec(d, box, k1, k2, theta, reuse_shell_sums=reuse)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            correlations.bispec(d, box, k1, k2, theta, reuse_shell_sums=reuse)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    ms_all = timed(False)      # every shell's indicator transformed (a first call)
    ms = timed(True)           # indicator sums reused (every later call with these shells)
    half = 8 * n * n * (n // 2 + 1)
    mesh = 4 * n ** 3
    _, _, _, hi = correlations.bispec_shells(n, box, k1, k2, theta)
    nsh = hi.size
    # with reused indicator sums (a shell touches the planes kz <= isqrt(hi - 1) only):
    # forward transform of the field up to the largest shell's plane (mesh read, z write,
    # y and x read + write); the first two shells' delta fields (x write, y read + write,
    # C2R read, mesh write); their product (two meshes read, one written); its forward
    # transform up to the later shells' largest plane; the Fourier-space sums (delta_k and
    # the product spectrum read over those planes, once per chunk of 8 shells)
    planes = [min(math.isqrt(int(h) - 1), n // 2) + 1 if h > 0 else 0 for h in hi]
    plane = 8 * n * n
    chunks = (nsh - 2 + 7) // 8
    pmax, plate = max(planes), max(planes[2:], default=0)
    byts = ((mesh + 5 * plane * pmax) + sum(4 * plane * p + mesh for p in planes[:2]) + 3 * mesh
            + (mesh + 5 * plane * plate) + chunks * 2 * plane * plate)
    gbs = byts / (ms * 1e-3) / 1e9
    out = {"workload": f"bispec {n}^3, L={box:g}, k1={k1}, k2={k2}, {theta.size} angles ({nsh} shells)",
           "ms": round(ms, 2), "per_s": round(1e3 / ms, 2), "ms_first_call": round(ms_all, 2),
           "algorithmic_gbs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
           "note": "ms: shell-indicator sums reused (they depend on n and the shells only); "
                   "ms_first_call: every indicator transformed"}
    del d
    torch.cuda.empty_cache()
    if cpu:
        from oracle import ref_fp64 as R
        ns = 128
        dh = np.random.default_rng(0).standard_normal((ns, ns, ns)).astype(np.float32)
        t0 = time.perf_counter()
        R.bispec(dh, box * ns / n, k1, k2, theta)
        dt = time.perf_counter() - t0
        scaled = dt * (n / ns) ** 3 * math.log2(n) / math.log2(ns)
        out["cpu_baseline"] = {"ms_scaled": round(scaled * 1e3, 1), "unit": "ms per bispec", "cores": 1,
                               "kind": "port", "extrapolated": True,
                               "sample": f"oracle FP64 NumPy bispec at {ns}^3 ({dt:.2f} s), scaled by n^3 log n "
                                         f"to {n}^3"}
        out["gpu_over_cpu"] = round(scaled * 1e3 / ms, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-timing", action="store_true", help="no per-stage events (for profiler runs)")
    ap.add_argument("--tile", type=str, default=None, help="1,ty,0 | direct")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device-to-device copy first (HBM counter calibration)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-mock batched comparison")
    ap.add_argument("--only-batched", action="store_true", help="run only the batched comparison")
    ap.add_argument("--no-served", action="store_true", help="skip the serving benchmark")
    ap.add_argument("--no-small", action="store_true", help="skip the 128^3 / 1e6-particle plain case")
    ap.add_argument("--only-served", action="store_true", help="run only the serving benchmark")
    ap.add_argument("--no-mock", action="store_true", help="skip the mock-generator measurement")
    ap.add_argument("--no-xi", action="store_true", help="skip the configuration-space measurement")
    ap.add_argument("--no-bispec", action="store_true", help="skip the bispectrum measurement")
    ap.add_argument("--only-bispec", action="store_true", help="run only the bispectrum measurement")
    ap.add_argument("--only-xi", action="store_true", help="run only the configuration-space measurement")
    ap.add_argument("--only-mock", action="store_true", help="run only the mock-generator measurement")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the calibrated HBM traffic summary of two rocprofv3 --pmc runs and exit")
    ap.add_argument("--pmc-command", default="", help="command line recorded in the --pmc-summary header")
    ap.add_argument("--trace-summary", nargs=3, metavar=("STATS_CSV", "OUT", "TITLE"),
                    help="write the library's kernels of a rocprofv3 kernel_stats.csv to OUT and exit")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "hbm_traffic_cic512_2e7.txt"),
                    help="traffic summary read for roofline.traffic")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.traffic, args.pmc_command)
        return
    if args.trace_summary:
        trace_summary(*args.trace_summary, args.pmc_command)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the path's "
              "multi-GPU sharding is a design, not a measured configuration", file=sys.stderr)
        sys.exit(2)

    import torch
    from jax_powspec_amd import _lib
    from jax_powspec_amd.correlations import Estimator

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.calib_copy:   # known-size copy: 1 GiB read + 1 GiB written per launch
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only_batched:
        print(json.dumps({"batched": bench_batched(dev)}), flush=True)
        return
    if args.only_served:
        print(json.dumps({"served": bench_served(dev)}), flush=True)
        return
    if args.only_bispec:
        print(json.dumps({"bispec": bench_bispec(dev, cpu=not args.no_cpu)}), flush=True)
        return
    if args.only_xi:
        print(json.dumps({"xi": bench_xi(dev, cpu=not args.no_cpu)}), flush=True)
        return
    if args.only_mock:
        print(json.dumps({"mock": bench_mock(dev, cpu=not args.no_cpu)}), flush=True)
        return
    log("generating catalogue ...")
    pos = make_catalogue(N_MESH, BOX, NBAR, dev)
    npart = pos.shape[0]
    log(f"N_p = {npart}")
    kf = 2 * np.pi / BOX
    edges = kf * (0.5 + np.arange(NBINS + 1))
    shot = BOX ** 3 / npart
    tile = None
    if args.tile:
        tile = "direct" if args.tile == "direct" else tuple(int(v) for v in args.tile.split(","))
        log(f"tile {tile}")
    est = Estimator(box_size=BOX, n_bins=N_MESH, max_particles=npart, k_edges=edges, mas="CIC_REF",
                    shot_noise=shot, device=dev, tile=tile)
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    for _ in range(args.warmup):
        est.run(x, y, z)
    torch.cuda.synchronize()
    lib = _lib.load()
    timing = not args.no_timing
    import ctypes as C

    def read_stages():
        msv = (C.c_double * _lib.NUM_STAGES)()
        nl = (C.c_int64 * _lib.NUM_STAGES)()
        _lib.check("pk_timing_read", lib.pk_timing_read(C.cast(msv, C.c_void_p), C.cast(nl, C.c_void_p),
                                                        _lib.NUM_STAGES))
        return {lib.pk_stage_name(i).decode(): {"avg_ms": msv[i] / nl[i], "launches": int(nl[i])}
                for i in range(_lib.NUM_STAGES) if nl[i]}

    stages = {}
    dom_stage = None
    if timing:
        # every stage's HIP events over a calibration pass of the same length (outside the
        # timed region: per-stage events add gaps between the launches); then only the
        # dominant stage carries events inside the timed region
        lib.pk_timing_read(None, None, 0)
        _lib.check("pk_timing_select", lib.pk_timing_select(-1))
        lib.pk_timing_enable(1)
        for _ in range(args.steps):
            est.run(x, y, z)
        torch.cuda.synchronize()
        lib.pk_timing_enable(0)
        stages = read_stages()
        for v in stages.values():
            v["pass"] = "calibration"
        dom_stage = max(stages, key=lambda s: stages[s]["avg_ms"])
        names = [lib.pk_stage_name(i).decode() for i in range(_lib.NUM_STAGES)]
        _lib.check("pk_timing_select", lib.pk_timing_select(names.index(dom_stage)))
        lib.pk_timing_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        est.run(x, y, z)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    if timing:
        lib.pk_timing_enable(0)
        live = read_stages()
        _lib.check("pk_timing_select", lib.pk_timing_select(-1))
        if dom_stage in live:   # the dominant kernel, timed live in the timed region
            stages[dom_stage] = dict(live[dom_stage], **{"pass": "timed"})
    pk = est.out.pk.cpu().numpy()
    gpu_sums = est.out.sums.cpu().numpy()
    gpu_counts = est.out.counts.cpu().numpy()
    log(f"P_0 first bins: {pk[:4, 0]}")
    roof = None
    if stages:
        fused = "zpass" not in stages
        ab = algorithmic_bytes(N_MESH, npart, fused)
        dom = max(stages, key=lambda s: stages[s]["avg_ms"])
        for s, v in stages.items():
            if s in ab:
                v["gbs"] = ab[s] / (v["avg_ms"] * 1e-3) / 1e9
        log("stages: " + json.dumps(stages))
        if dom in ab:
            ach = ab[dom] / (stages[dom]["avg_ms"] * 1e-3) / 1e9
            kname = stage_kernel(dom, N_MESH, fused)
            traffic = None
            for k, (f, w) in read_traffic(args.traffic).items():
                if kname in k:
                    traffic = round(f + w)
            roof = {"bound": "hbm", "kernel": kname, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "algorithmic_bytes": ab[dom], "avg_ms": round(stages[dom]["avg_ms"], 5)}
        path_bytes = 12 * npart + 2 * 4 * N_MESH ** 3 + 4 * 8 * N_MESH * N_MESH * (N_MESH // 2 + 1)
        roof_path = path_bytes / (ms * 1e-3) / 1e9
        log(f"whole-path algorithmic bytes {path_bytes / 1e9:.3f} GB -> {roof_path:.0f} GB/s "
            f"({roof_path / HBM_PEAK_GBS:.3f} of peak)")
    batched = served = None
    del est
    small = None if args.no_small else bench_small(dev, cpu=not args.no_cpu)
    log(f"small: {small}")
    if not args.no_batched:
        batched = bench_batched(dev)
        torch.cuda.empty_cache()
    if not args.no_served:
        served = bench_served(dev)
        torch.cuda.empty_cache()
    mock = None
    if not args.no_mock:
        mock = bench_mock(dev, cpu=not args.no_cpu)
        log(f"mock: {mock}")
        torch.cuda.empty_cache()
    xi = None
    if not args.no_xi:
        xi = bench_xi(dev, cpu=not args.no_cpu)
        log(f"xi: {xi}")
        torch.cuda.empty_cache()
    bis = None
    if not args.no_bispec:
        bis = bench_bispec(dev, cpu=not args.no_cpu)
        log(f"bispec: {bis}")
        torch.cuda.empty_cache()
    cpu = parity = None
    if not args.no_cpu:
        log("cpu baseline ...")
        pos_h = pos.cpu().numpy()
        cpu = cpu_baseline(pos_h, N_MESH, BOX, edges, shot)
        log(f"cpu baseline: {cpu}")
        log("parity against the FP64 oracle ...")
        parity = parity_vs_oracle(pos_h, N_MESH, BOX, edges, shot, gpu_sums, gpu_counts)
        log(f"parity: {parity}")
    out = {
        "metric": "P_l(k) estimates/sec at 512^3, 2e7 particles; batched/served vs plain; p99 ms",
        "value": round(1e3 / ms, 3),
        "unit": "estimates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 per-bin sums)",
        "data": "synthetic",
        "config": {"workload": f"lognormal+RSD mock, L=1500 Mpc/h, 512^3 CIC (reference stencil), N_p={npart}, "
                               "P_0,2,4 in 256 k-bins", "mesh": N_MESH, "particles": npart, "bins": NBINS},
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "small": small,
        "batched": batched,
        "served": served,
        "mock": mock,
        "xi": xi,
        "bispec": bis,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
