#!/usr/bin/env python3
"""Batched 2-D CG benchmark on one MI355X.

Workload (the batched 2-D CG configuration): 512 independent slices of uneven
size, N drawn uniformly from {128, 192, 256, 384, 512, 768, 1024}, angle counts
from {90, 180, 360, 720} over [0, 180), detector width N; float32
random-ellipse phantoms (10 ellipses), sinograms from the NUFFT forward
projector with Gaussian noise of 1% of the maximum, normalised by the maximum
(include/mbir2d_prep.h:45-65); right-hand side = backproj2d; CG with the
Toeplitz normal operator, 50 iterations, tol 1e-5, xtol 1e-5 (the recon_tomo
defaults, include/config.h:365-385).  Seeded synthetic data.

One step = one batched solve of all slices (every slice its own problem with
its own CG scalars), replayed from a captured hipGraph, inputs resident in
HBM.  Reported beside it:
  * the plain path: the same slices solved one at a time through lc_cg2d
    (host arrays in and out, host-synchronised iterations), on a sample;
  * the same batched solve in the FP64 path (float64 vectors, spectra, FFTs);
  * relative L2 error against the FP64 oracle on a sample of slices, for the
    FP32 and the FP64 path, beside independent FP64 realisations;
  * roofline of the dominant Toeplitz pass from HIP events on the library's
    stream, right after the timed steps, on the same batch;
  * the FP32 CPU restatement of the reference path (oracle, OpenMP) timed on
    the host on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N_SET = (128, 192, 256, 384, 512, 768, 1024)
ANG_SET = (90, 180, 360, 720)
HBM_PEAK_GBS = 8000.0
PASS_NAMES = ("rows_fwd", "cols", "rows_inv", "cg_update")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(nslices, seed_stream=4 << 16):
    from laminocam_amd import synth
    rng = synth.Stream(seed_stream)
    Ns = rng.choice(N_SET, nslices).astype(int)
    As = rng.choice(ANG_SET, nslices).astype(int)
    return rng, Ns, As


def build_problem(lc, synth, rng, Ns, As):
    """geometries, phantoms, noisy normalised sinograms and right-hand sides"""
    geoms = {}
    for N, a in zip(Ns, As):
        key = (int(N), int(a))
        if key not in geoms:
            grid = lc.PolarGrid2D(synth.angles(int(a)), int(N))
            geoms[key] = (grid, lc.PointSpreadFunction2D(grid, int(N)))
    groups = {}
    for i, (N, a) in enumerate(zip(Ns, As)):
        groups.setdefault((int(N), int(a)), []).append(i)
    rhs = [None] * len(Ns)
    for key, idx in sorted(groups.items()):
        grid, _ = geoms[key]
        N = key[0]
        vol = np.stack([synth.ellipse_phantom(N, 10, synth.Stream((4 << 16) + 1 + i)) for i in idx])
        sino = lc.forward2d(vol, grid)
        for j, i in enumerate(idx):
            s = synth.add_gaussian_noise(sino[j], 0.01, synth.Stream((4 << 16) + 100000 + i))
            sino[j] = s / np.float32(np.max(np.abs(s)))
        y = lc.backproj2d(sino, grid, (len(idx), N, N))
        for j, i in enumerate(idx):
            rhs[i] = y[j:j + 1]
    return geoms, rhs


ERR_HORIZONS = (5, 8, 10, 15, 20, 30, 40)


def rel_err_by_class(args, lc, synth, Ns, As, psfs, rhs, xb):
    """Relative L2 error against the FP64 oracle on the first slice of every
    (N, angle count) class of the workload, at the horizons 5, 8, 10, ... up to
    the workload's iteration count, for the GPU's FP32 path, the GPU's FP64
    path (an FP64 realisation whose roundings differ from the oracle's: other
    FFT algorithm, other summation order) and the FP32 restatement of the
    reference's own CPU path (oracle in float at the reference's float NUFFT
    tolerance, full horizon only).  The oracle rerun on a right-hand side
    perturbed by 1e-15 relative is reported too; it repeats the oracle's
    roundings, so it stays closer than any independent realisation.  Per
    class: the longest horizon at which the GPU's FP32 result is within 1e-3
    of the oracle, and the longest at which the FP64 realisation is (past it,
    rounding alone moves a result by more than 1e-3).  The (256, 180) class is
    the plain 2-D configuration (256^2 ellipse phantom, 180 angles, 256 bins,
    CG 50 iterations) and is also reported on its own, with its plain GPU and
    CPU solve times."""
    from oracle.oracle import Oracle
    o64, o32 = Oracle(64), Oracle(32)
    H = [k for k in ERR_HORIZONS if k < args.iters] + [args.iters]
    first = {}
    for i, (N, a) in enumerate(zip(Ns, As)):
        first.setdefault((int(N), int(a)), i)
    rows, cfg1 = [], None
    t_start = time.perf_counter()
    r3 = lambda v: float("%.3g" % v)
    for (N, a), i in sorted(first.items()):
        th = synth.angles(a)
        y = rhs[i]
        y64 = y.astype(np.float64)
        z = np.zeros(y.shape)
        P = o64.psf2d(th, N, N, method=1)
        xo, ro = o64.cgsolver_trace(P, y64, z, H, args.tol, args.xtol)
        yp = y64 * (1.0 + 1e-15 * synth.Stream((4 << 16) + 200000 + i).normal(y.size)
                    .reshape(y.shape))
        xp, _ = o64.cgsolver_trace(P, yp, z, H, args.tol, args.xtol)
        del P
        e32, e64, fl = [], [], []
        for j, k in enumerate(H):
            g32 = xb[i] if k == args.iters else \
                lc.cgsolver(psfs[i], y, np.zeros_like(y), k, args.tol, args.xtol)[0]
            g64 = lc.cgsolver(psfs[i], y64, z, k, args.tol, args.xtol)[0]
            e32.append(rel_l2(g32, xo[j]))
            e64.append(rel_l2(g64, xo[j]))
            fl.append(rel_l2(xp[j], xo[j]))
        agree = [k for j, k in enumerate(H) if e64[j] <= 1e-3]
        hz = max(agree) if agree else None
        ok32 = [k for j, k in enumerate(H) if e32[j] <= 1e-3]
        P32 = o32.psf2d(th, N, N, method=1)
        t0 = time.perf_counter()
        x32, _ = o32.cgsolver(P32, y, np.zeros_like(y), args.iters, args.tol, args.xtol)
        t_port = time.perf_counter() - t0
        del P32
        row = {"N": N, "nang": a, "fp32": r3(e32[-1]), "fp64_gpu": r3(e64[-1]),
               "fp64_rerun": r3(fl[-1]), "fp32_port": r3(rel_l2(x32, xo[-1])),
               "fp32_horizon_1e-3": max(ok32) if ok32 else None,
               "fp64_horizon_1e-3": hz, "fp32_at_fp64_horizon": r3(e32[H.index(hz)]) if hz else None,
               "oracle_iters": int(ro["iters"]),
               "by_iters": {"fp32": [r3(e) for e in e32], "fp64_gpu": [r3(e) for e in e64],
                            "fp64_rerun": [r3(e) for e in fl]}}
        rows.append(row)
        if (N, a) == (256, 180):
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                lc.cgsolver(psfs[i], y, np.zeros_like(y), args.iters, args.tol, args.xtol)
                ts.append(time.perf_counter() - t0)
            o32t = int(o32.lib.or_num_threads())
            cfg1 = {"workload": f"plain 2-D CG: 256^2 ellipse phantom (10 ellipses), 180 angles "
                                f"over [0,180), 256 bins, 1% Gaussian noise, {args.iters} "
                                f"iterations, tol {args.tol:g}",
                    "gpu_plain_ms": round(1e3 * float(np.median(ts)), 3),
                    "gpu_plain_slices_per_s": round(1.0 / float(np.median(ts)), 2),
                    "cpu_port_s": round(t_port, 3), "cpu_threads": o32t,
                    "rel_err_vs_fp64": {k: row[k] for k in ("fp32", "fp64_gpu", "fp64_rerun",
                                                            "fp32_port", "fp32_horizon_1e-3",
                                                            "fp64_horizon_1e-3")},
                    "note": "gpu_plain: lc_cg2d, host arrays in and out, host-synchronised "
                            "iterations; cpu_port: the FP32 C restatement of the reference "
                            "path (PSF excluded)"}
        log(f"err class N={N} nang={a}: fp32 {e32[-1]:.2e} fp64 gpu {e64[-1]:.2e} rerun "
            f"{fl[-1]:.2e} port {row['fp32_port']:.2e}; within 1e-3 up to {row['fp32_horizon_1e-3']} "
            f"it (FP32), {hz} it (FP64)")
    h32 = [r["fp32_horizon_1e-3"] or 0 for r in rows]
    h64 = [r["fp64_horizon_1e-3"] or 0 for r in rows]
    summary = {
        "fp32_within_1e-3_at_full_horizon": sum(r["fp32"] <= 1e-3 for r in rows),
        "fp64_gpu_within_1e-3_at_full_horizon": sum(r["fp64_gpu"] <= 1e-3 for r in rows),
        "fp32_horizon_1e-3_min_max": [min(h32), max(h32)],
        "fp64_horizon_1e-3_min_max": [min(h64), max(h64)],
        "fp32_over_reference_float_path_median": r3(float(np.median(
            [r["fp32"] / max(r["fp32_port"], 1e-30) for r in rows]))),
        "seconds": round(time.perf_counter() - t_start, 1)}
    return {"iters": args.iters, "horizons": H, "classes": rows, "summary": summary,
            "config1": cfg1,
            "note": "per class (first slice of each N x angles class): fp32 = GPU FP32 path "
                    "(the batched result), fp64_gpu = GPU FP64 path, fp64_rerun = FP64 oracle on "
                    "a right-hand side perturbed by 1e-15 (same roundings, so not a floor), "
                    "fp32_port = FP32 restatement of the reference CPU path, all against the FP64 "
                    "oracle at the full horizon; fp32_horizon_1e-3 / fp64_horizon_1e-3 = last "
                    "iteration count of 'horizons' at which that path is within 1e-3 of the "
                    "oracle; by_iters at every horizon"}


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cpu_baseline(args, iters_by_slice, Ns, rhs_by_n):
    """FP32 CPU restatement of the reference path (Toeplitz CG): per image
    size N one slice of the workload is solved cpu_reps times with cpu_iters
    CG iterations (the median is kept).  A Toeplitz CG iteration costs the same
    whatever the angle count, so the time of the whole workload is the sum
    over its slices of (seconds per iteration at that N) x (iterations the
    batched solve ran)."""
    from oracle.oracle import Oracle
    o32 = Oracle(32)
    threads = int(o32.lib.or_num_threads())
    t_iter = {}
    spent = 0.0
    for N in sorted(rhs_by_n):
        y, theta = rhs_by_n[N]
        psf = o32.psf2d(theta, N, N, method=0)
        k = args.cpu_iters
        ts = []
        for _ in range(args.cpu_reps):
            t0 = time.perf_counter()
            o32.cgsolver(psf, y, np.zeros_like(y), k, 0.0, 0.0)
            ts.append(time.perf_counter() - t0)
        spent += sum(ts)
        t_iter[N] = float(np.median(ts)) / k
    total = sum(t_iter[int(N)] * iters_by_slice[i] for i, N in enumerate(Ns))
    return {
        "value": len(Ns) / total,
        "unit": "slices/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP32 C restatement of the reference CPU path (Toeplitz CG), OpenMP "
                   f"{threads} threads: one workload slice per image size N in {sorted(t_iter)}, "
                   f"{args.cpu_iters} CG iterations, median of {args.cpu_reps} solves "
                   f"({spent:.1f} s of CPU solves); workload time = sum over the {len(Ns)} "
                   f"slices of (seconds per iteration at its N) x (iterations the batched "
                   f"solve ran)"),
        "sec_per_iter": {str(k): v for k, v in t_iter.items()},
    }


def served_stream(args, lc, synth, ctx, frac3d=None, rate=None, sweep=None):
    """Served mixed stream: Poisson arrivals at a fixed rate of 2-D slices
    (N 128-1024, 180-720 angles) and 3-D laminography volumes (8-64 slices x
    256-511^2, 180-720 projections, 30 degree tilt), 70% CG / 30% Split-Bregman
    TV (lambda U[0.05, 0.2], mu 10, zero-padded differences, so that the
    shrink acts on single slices too), submitted as sinograms / projections to
    the reconstruction server (HBM-budget admission, continuous batching).
    Inputs come from a pool of seeded phantoms per geometry built before the
    clock.  frac3d / rate override --serve-3d-frac / --serve-rate; sweep is a
    list of arrival rates run for --serve-sweep-seconds each, whose highest
    rate the server keeps up with (90% of the offered requests per second over
    the span from the first submission to the last result, p99 latency under
    1 s) is reported as the saturating rate."""
    frac3d = args.serve_3d_frac if frac3d is None else frac3d
    rate = args.serve_rate if rate is None else rate
    import threading
    geo2, pool2 = [], []
    for N in N_SET:
        for a in (180, 360, 720):
            grid = lc.PolarGrid2D(synth.angles(a), N)
            psf = lc.PointSpreadFunction2D(grid, N, ctx=ctx)
            ph = synth.ellipse_phantom(N, 10, synth.Stream((5 << 16) + 100 + len(geo2)))[None]
            s = lc.forward2d(ph, grid, ctx=ctx)
            s = synth.add_gaussian_noise(s, 0.01, synth.Stream((5 << 16) + 200 + len(geo2)))
            geo2.append(psf)
            pool2.append(s / np.float32(np.max(np.abs(s))))
    vols = [((8, 256, 256), 180, 64), ((16, 384, 384), 360, 96), ((32, 511, 511), 720, 128),
            ((64, 511, 511), 720, 192)]
    if frac3d <= 0.0:
        vols = []
    geo3, pool3 = [], []
    for i, (dims, nproj, nrows) in enumerate(vols):
        th = np.arange(nproj) * (2 * np.pi / nproj)
        grid = lc.PolarGrid(th, np.full(nproj, -np.pi / 6), np.zeros(nproj), nrows, dims[2])
        psf = lc.PointSpreadFunction(grid, dims, ctx=ctx)
        ph = synth.ellipsoid_phantom(dims, 12, synth.Stream((5 << 16) + 300 + i))
        p = lc.forward(ph, grid, ctx=ctx)
        p = synth.add_gaussian_noise(p, 0.01, synth.Stream((5 << 16) + 400 + i))
        geo3.append(psf)
        pool3.append(p / np.float32(np.max(np.abs(p))))

    def make_requests(n, seed):
        rng = synth.Stream(seed)
        kinds, solv = rng.uniform(n), rng.uniform(n)
        picks2 = rng.choice(list(range(len(geo2))), n)
        picks3 = rng.choice(list(range(max(1, len(geo3)))), n)
        lams = rng.uniform(n, 0.05, 0.2)
        reqs = []
        for i in range(n):
            is3 = kinds[i] < frac3d
            psf = geo3[picks3[i]] if is3 else geo2[picks2[i]]
            sino = pool3[picks3[i]] if is3 else pool2[picks2[i]]
            sb = solv[i] < 0.3
            slices = psf.recon_dims[0] if is3 else 1
            kw = dict(solver="sb", lam=float(lams[i]), mu=10.0, outer=args.sb_outer, inner=1,
                      zero_pad=True) if sb else dict(solver="cg", max_iter=args.iters)
            reqs.append((psf, sino, kw, slices, is3, sb))
        return reqs

    def run(srv, reqs, r):
        """Poisson arrivals at r requests/s from a submitting thread; every
        result taken; latency = submit -> result ready"""
        gaps = -np.log(1.0 - synth.Stream((5 << 16) + 7 + int(r)).uniform(len(reqs))) / r
        tickets = [None] * len(reqs)
        t0 = time.perf_counter()
        arrive = t0 + np.cumsum(gaps)

        def submitter():
            for i, (psf, sino, kw, _, _, _) in enumerate(reqs):
                d = arrive[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(psf, sino=sino, tol=args.tol, xtol=args.xtol, borrow=True,
                                        **kw)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        tm = [srv.result(t, timeout=600.0)[2] for t in tickets]
        lat = np.array([t[2] - t[0] for t in tm])
        span = max(t[2] for t in tm) - min(t[0] for t in tm)
        slices = sum(q[3] for q in reqs)
        return {"rate_req_per_s": r, "requests": len(reqs), "slices": int(slices),
                "slices_per_s": round(slices / span, 2), "requests_per_s": round(len(reqs) / span, 2),
                "latency_ms": {"p50": round(1e3 * float(np.median(lat)), 2),
                               "p99": round(1e3 * float(np.percentile(lat, 99)), 2)}}

    n = max(1, int(round(rate * args.serve_seconds)))
    reqs = make_requests(n, (5 << 16) + 1)
    with lc.Server(hbm_budget=int(args.serve_budget_gb * (1 << 30)), lanes=args.serve_lanes,
                   max_batch=args.serve_max_batch, ctx=ctx) as srv:
        # warm the lanes' kernels and scratch on one request of each kind
        for psf, sino in [(geo2[0], pool2[0])] + ([(geo3[0], pool3[0])] if geo3 else []):
            srv.result(srv.submit(psf, sino=sino, max_iter=2), timeout=120.0)
        st0 = srv.stats()
        main = run(srv, reqs, rate)
        st = srv.stats()
        # rate sweep
        sweep_out, saturating = [], None
        for r in (sweep or []):
            m = max(1, int(round(r * args.serve_sweep_seconds)))
            rec = run(srv, make_requests(m, (5 << 16) + 1000 + int(r)), r)
            rec["keeps_up"] = bool(rec["requests_per_s"] >= 0.9 * r and
                                   rec["latency_ms"]["p99"] < 1000.0)
            sweep_out.append(rec)
            log(f"served sweep {r:g} req/s: {rec['slices_per_s']} slices/s, p50 "
                f"{rec['latency_ms']['p50']} ms p99 {rec['latency_ms']['p99']} ms"
                f"{'' if rec['keeps_up'] else ' (falls behind)'}")
            if rec["keeps_up"]:
                saturating = rec
        # capacity: the same requests submitted at once, timed to the last result
        st_c0 = srv.stats()
        t_b0 = time.perf_counter()
        tk = [srv.submit(psf, sino=sino, tol=args.tol, xtol=args.xtol, borrow=True, **kw)
              for psf, sino, kw, _, _, _ in reqs]
        burst = [srv.result(t, timeout=600.0)[2] for t in tk]
        t_b = time.perf_counter() - t_b0
        st_b = srv.stats()
        tl = srv.timeline()
    slices = sum(r[3] for r in reqs)
    # the lanes' device intervals during the capacity run: time with two or
    # more lanes' solves in flight at once
    cap_tl = tl[tl[:, 2] >= min(b[1] for b in burst) - 1e-3] if len(tl) else tl
    overlap_s = 0.0
    if len(cap_tl):
        ev = sorted([(r[4], 1) for r in cap_tl] + [(r[5], -1) for r in cap_tl])
        depth, last = 0, None
        for t, d in ev:
            if depth >= 2 and last is not None:
                overlap_s += (t - last) / 1e3
            depth += d
            last = t
    # the plain path on the same mix: the first requests of the stream solved
    # one at a time through the single-problem entry points (back-projection,
    # then cgsolver / split_bregman, host arrays in and out)
    k = min(n, args.serve_plain_sample)
    t0 = time.perf_counter()
    for psf, sino, kw, _, is3, sb in reqs[:k]:
        if is3:
            rhs = lc.backproj(sino, psf.grid, psf.recon_dims, ctx=ctx)
        else:
            N = psf.image_size
            rhs = lc.backproj2d(sino, psf.grid, (sino.shape[0], N, N), ctx=ctx)
        x0 = np.zeros(rhs.shape, np.float32)
        if sb:
            lc.split_bregman(psf, rhs, x0, kw["lam"], kw["mu"], kw["outer"], kw["inner"], args.tol,
                             args.xtol, zero_pad=True)
        else:
            lc.cgsolver(psf, rhs, x0, kw["max_iter"], args.tol, args.xtol)
    plain_rate = sum(r[3] for r in reqs[:k]) / (time.perf_counter() - t0)
    out = dict(main)
    out.update({
        "seconds_of_arrivals": args.serve_seconds,
        "mix": {"3d": int(sum(r[4] for r in reqs)), "split_bregman": int(sum(r[5] for r in reqs))},
        "plain_same_mix": {"slices_per_s": round(plain_rate, 2), "sample_requests": k},
        "speedup_vs_plain_same_mix": round(main["slices_per_s"] / plain_rate, 2),
        "capacity": {"slices_per_s": round(slices / t_b, 2), "seconds": round(t_b, 3),
                     "speedup_vs_plain_same_mix": round(slices / t_b / plain_rate, 2),
                     "latency_ms_p50": round(1e3 * float(np.median([b[2] - b[0] for b in burst])), 2),
                     "batches": int(st_b["batches"] - st_c0["batches"]),
                     "lane_seconds": {k[7:]: round(st_b[k] - v, 3) for k, v in st_c0.items()
                                      if k.startswith("lane_s_")},
                     "device_seconds_with_2_or_more_lanes_solving": round(overlap_s, 3),
                     "last_submit_s": round(max(b[0] for b in burst) - min(b[0] for b in burst), 3),
                     "note": "every request of the stream submitted at once"},
        "sweep": sweep_out,
        "saturating": saturating,
        "lanes": args.serve_lanes, "max_batch": args.serve_max_batch, "budget_gb": args.serve_budget_gb,
        "batches": int(st["batches"] - st0["batches"]),
        "lane_seconds": {k[7:]: round(v - st0[k], 3) for k, v in st.items() if k.startswith("lane_s_")},
        "workload": f"Poisson arrivals; 2-D slices N 128-1024 x 180/360/720 angles and 3-D volumes "
                    f"8x256^2 / 16x384^2 / 32x511^2 / 64x511^2 (180/360/720/720 projections, 30 "
                    f"deg tilt) as {100 * frac3d:g}% of requests; 70% CG "
                    f"({args.iters} it), 30% Split-Bregman ({args.sb_outer} outer x 1 inner, "
                    "lambda U[0.05,0.2], mu 10, zero-padded differences); sinograms "
                    "back-projected on the device; host arrays in and out (inputs lent to the "
                    "server, solutions written into the caller's arrays)",
    })
    return out


def lamino_leg(args, lc, synth, ctx):
    """3-D laminography solves at the padded shapes of the laminography
    configurations (src/mbir.cpp:34-161 pads recon [21, 511, 511] to
    [29, 721, 721] and projections 360 x 64 x 511 to 360 x 89 x 721):
    360 angles over [0, 360) at 30 degree tilt, PSF (64, 1458, 1458) in FP64.
      CG:  12-ellipsoid phantom, 1% Gaussian noise, 100 iterations.
      SB:  piecewise-constant phantom, Poisson counts at 1e4 photons / ray
           (I = Poisson(1e4 exp(-s p)), s = 2 / max p, y = -ln(max(I, 1) / 1e4) / s),
           lambda 0.1, mu 10, inner 1, 100 outer iterations, isotropic shrink,
           interior-only differences.
    Projections come from the GPU forward projector; timing covers the solve
    only (inputs resident)."""
    dims = (29, 721, 721)
    nproj = 360
    th = np.arange(nproj) * (2 * np.pi / nproj)
    grid = lc.PolarGrid(th, np.full(nproj, -np.pi / 6), np.zeros(nproj), 89, 721)
    t0 = time.perf_counter()
    psf = lc.PointSpreadFunction(grid, dims, ctx=ctx)
    ctx.sync()
    t_psf = time.perf_counter() - t0
    out = {"dims": list(dims), "projections": [nproj, 89, 721], "psf_dims": list(psf.dims()),
           "psf_build_s": round(t_psf, 3)}
    # CG (configuration 2)
    ph = synth.ellipsoid_phantom(dims, 12, synth.Stream((2 << 16) + 1))
    p = lc.forward(ph, grid, ctx=ctx)
    p = synth.add_gaussian_noise(p, 0.01, synth.Stream((2 << 16) + 2))
    p /= np.float32(np.max(np.abs(p)))
    t0 = time.perf_counter()
    y = lc.backproj(p, grid, dims, ctx=ctx)
    ctx.sync()
    t_bp = time.perf_counter() - t0
    b = lc.Batch2D([psf], ctx=ctx)
    b.set(0, b.Y, y)
    b.cg(2, args.tol, args.xtol, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    b.cg(args.lamino_iters, 0.0, 0.0, graph=True)
    ctx.sync()
    t_cg = time.perf_counter() - t0
    rep = b.reports()[0]
    out["cg"] = {"iters": int(rep["iters"]), "seconds": round(t_cg, 4),
                 "ms_per_iter": round(1e3 * t_cg / max(1, rep["iters"]), 3),
                 "backproj_s": round(t_bp, 4), "res": rep["res"]}
    # Split-Bregman (configuration 3)
    ph = synth.ellipsoid_phantom(dims, 12, synth.Stream((3 << 16) + 1), piecewise=True)
    pr = lc.forward(ph, grid, ctx=ctx).astype(np.float64)
    s = 2.0 / max(float(np.max(pr)), 1e-12)
    counts = np.random.default_rng((3 << 16) + 2).poisson(1e4 * np.exp(-s * np.maximum(pr, 0.0)))
    yp = (-np.log(np.maximum(counts, 1) / 1e4) / s).astype(np.float32)
    y = lc.backproj(yp, grid, dims, ctx=ctx)
    b.set(0, b.Y, y)
    b.split_bregman(0.1, 10.0, 2, 1, 0.0, 0.0)
    ctx.sync()
    t0 = time.perf_counter()
    b.split_bregman(0.1, 10.0, args.lamino_iters, 1, 0.0, 0.0)
    ctx.sync()
    t_sb = time.perf_counter() - t0
    rep = b.reports()[0]
    out["split_bregman"] = {"outer": int(rep["outer"]), "seconds": round(t_sb, 4),
                            "ms_per_outer": round(1e3 * t_sb / max(1, rep["outer"]), 3)}
    # slices/s: the reference counts the cropped volume's recon_dims[0] = 21
    # slices (src/mbir.cpp:159 crops 29 -> 21)
    out["cg"]["slices_per_s"] = round(21 / t_cg, 2)
    out["split_bregman"]["slices_per_s"] = round(21 / t_sb, 2)
    out["slices_note"] = "slices/s counts the 21 slices of the cropped reconstruction"
    if not args.no_err:
        out["accuracy_reduced_shape"] = lamino_accuracy(args, lc, synth, ctx)
    if not args.no_cpu and not args.no_lamino_cpu:
        t_apply, threads = lamino_cpu_apply_s(dims, tuple(psf.dims()))
        t_cpu = t_apply * (args.lamino_iters + 1)
        out["cpu_baseline"] = {
            "value": round(21 / t_cpu, 4), "unit": "slices/s", "cores": threads, "kind": "port",
            "seconds_per_apply": round(t_apply, 3),
            "sample": f"one Toeplitz apply of the FP32 C restatement at {list(dims)} (PSF "
                      f"{list(psf.dims())}), times {args.lamino_iters + 1} applies of a "
                      f"{args.lamino_iters}-iteration CG",
            "speedup_gpu_vs_cpu": round(t_cpu / t_cg, 1)}
    out["mbir_driver"] = lamino_driver(args, lc, synth, ctx)
    out["workload"] = ("3-D laminography at the padded configuration shapes; CG 100 iterations "
                       "(1% Gaussian noise) and Split-Bregman TV 100 outer x 1 inner (Poisson "
                       "counts, lambda 0.1, mu 10, isotropic, interior-only differences); "
                       "synthetic phantoms, inputs resident")
    return out


def lamino_driver(args, lc, synth, ctx):
    """The whole laminography reconstruction through the MBIR driver
    (lc_mbir3d, src/mbir.cpp:34-161), as recon_lamino runs it: projections
    360 x 64 x 511 over [0, 360) at 30 degree tilt, recon_dims 21 x 511^2,
    PAD_FACTOR 1.4142 (padded to 29 x 721^2 and 360 x 89 x 721), CG with
    args.lamino_iters iterations (tol, xtol 0).  Host arrays in and out; the
    driver builds the FP64 PSF, back-projects, solves and crops inside the
    timed call.  The second of two calls is kept."""
    from laminocam_amd import drivers
    nproj, nrows, ncols, rdims = 360, 64, 511, (21, 511, 511)
    th = np.arange(nproj) * (2 * np.pi / nproj)
    grid = lc.PolarGrid(th, np.full(nproj, -np.pi / 6), np.zeros(nproj), nrows, ncols)
    ph = synth.ellipsoid_phantom(rdims, 12, synth.Stream((2 << 16) + 21))
    p = lc.forward(ph, grid, ctx=ctx)
    p = synth.add_gaussian_noise(p, 0.01, synth.Stream((2 << 16) + 22))
    ds = [drivers.Dataset(projs=p, angles=th, gamma=-np.pi / 6)]
    par = drivers.ReconParams(recon_dims=rdims, maxIters=args.lamino_iters, tol=0.0, xtol=0.0)
    ts, rep = [], {}
    for _ in range(2):
        rep = {}
        t0 = time.perf_counter()
        x = drivers.MBIR(ds, par, ctx=ctx, graph=True, report=rep)
        ts.append(time.perf_counter() - t0)
    c = float(np.sum(x * ph) / (np.linalg.norm(x) * np.linalg.norm(ph)))
    return {"projections": [nproj, nrows, ncols], "recon_dims": list(rdims),
            "padded_dims": [int(v) for v in rep.get("padded_dims", [])],
            "iters": int(rep.get("iters", 0)), "seconds": round(ts[-1], 3),
            "slices_per_s": round(rdims[0] / ts[-1], 2),
            "cosine_to_phantom": round(c, 3),
            "note": "end to end through lc_mbir3d: PSF build, back-projection, CG, crop; host "
                    "arrays in and out"}


def lamino_accuracy(args, lc, synth, ctx):
    """Relative L2 error against the FP64 oracle of the 3-D CG (up to 100
    iterations) and Split-Bregman (20 outer x 1 inner, lambda 0.1, mu 10,
    isotropic, interior-only differences) solves on a reduced laminography
    shape the oracle can run: volume 9 x 96^2, 360 projections of 16 x 96 over
    [0, 360) at 30 degree tilt, 12-ellipsoid phantom, 1% noise.  Beside each:
    the GPU's FP64 path and the oracle rerun on a right-hand side perturbed by
    1e-15, the rounding floor of the problem."""
    from oracle.oracle import Oracle
    o64 = Oracle(64)
    dims, nproj, nrows = (9, 96, 96), 360, 16
    th = np.arange(nproj) * (2 * np.pi / nproj)
    ga, be = np.full(nproj, -np.pi / 6), np.zeros(nproj)
    grid = lc.PolarGrid(th, ga, be, nrows, dims[2])
    psf = lc.PointSpreadFunction(grid, dims, ctx=ctx)
    ph = synth.ellipsoid_phantom(dims, 12, synth.Stream((2 << 16) + 11))
    p = lc.forward(ph, grid, ctx=ctx)
    p = synth.add_gaussian_noise(p, 0.01, synth.Stream((2 << 16) + 12))
    p /= np.float32(np.max(np.abs(p)))
    y = lc.backproj(p, grid, dims, ctx=ctx)
    y64 = y.astype(np.float64)
    yp = y64 * (1.0 + 1e-15 * synth.Stream((2 << 16) + 13).normal(y.size).reshape(y.shape))
    P = o64.psf3d(th, ga, be, nrows, dims[2], dims, method=0)
    z = np.zeros(dims)
    r3 = lambda v: float("%.3g" % v)
    H = [5, 20, 50, args.lamino_iters]
    xo, _ = o64.cgsolver_trace(P, y64, z, H, 0.0, 0.0)
    xp, _ = o64.cgsolver_trace(P, yp, z, H, 0.0, 0.0)
    cg = {"iters": H, "fp32": [], "fp64_gpu": [], "fp64_rerun": []}
    for j, k in enumerate(H):
        cg["fp32"].append(r3(rel_l2(lc.cgsolver(psf, y, np.zeros_like(y), k, 0.0, 0.0)[0], xo[j])))
        cg["fp64_gpu"].append(r3(rel_l2(lc.cgsolver(psf, y64, z, k, 0.0, 0.0)[0], xo[j])))
        cg["fp64_rerun"].append(r3(rel_l2(xp[j], xo[j])))
    kw = (0.1, 10.0, 20, 1, 0.0, 0.0)
    so, _ = o64.split_bregman(P, y64, z, *kw)
    sp, _ = o64.split_bregman(P, yp, z, *kw)
    sb = {"outer": 20, "fp32": r3(rel_l2(lc.split_bregman(psf, y, np.zeros_like(y), *kw)[0], so)),
          "fp64_gpu": r3(rel_l2(lc.split_bregman(psf, y64, z, *kw)[0], so)),
          "fp64_rerun": r3(rel_l2(sp, so))}
    return {"shape": {"dims": list(dims), "projections": [nproj, nrows, dims[2]], "tilt_deg": 30},
            "cg": cg, "split_bregman": sb,
            "note": "against the FP64 oracle on the same right-hand side; fp64_gpu: the GPU's "
                    "FP64 path; fp64_rerun: the oracle on a right-hand side perturbed by 1e-15"}


def lamino_cpu_apply_s(dims, psf_dims):
    """Seconds of one application of the 3-D Toeplitz operator by the FP32 C
    restatement of the reference path (oracle in float, OpenMP) at the padded
    laminography shape: one complex FFT pair over the PSF dims and the
    pointwise product, the cost of one CG iteration.  The PSF values used for
    this timing are a centred unit impulse; the operation count does not
    depend on them.  The second of two applications is kept."""
    from oracle.oracle import Oracle
    o32 = Oracle(32)
    P = np.zeros(psf_dims, np.float32)
    P[tuple(d // 2 for d in psf_dims)] = 1.0
    Q = o32.psf_from_real(P)
    del P
    x = synth_random(dims)
    ts = []
    for _ in range(2):
        t0 = time.perf_counter()
        Q.convolve(x)
        ts.append(time.perf_counter() - t0)
    return ts[-1], int(o32.lib.or_num_threads())


def synth_random(dims):
    return np.random.default_rng(5).random(dims).astype(np.float32)


def nufft_probe(reps=2):
    """Back-projections for counter runs of the spreads: 8 slices of 1024^2
    from 720 angles x 1024 bins (2-D), and a 32 x 511^2 volume from 720
    projections of 128 x 511 at 30 degrees tilt (3-D); device-resident
    inputs, HIP-event times printed as one JSON line."""
    import ctypes as C
    import laminocam_amd as lc
    from laminocam_amd import synth
    ctx = lc.default_context()
    hip = C.CDLL("libamdhip64.so")
    out = {}

    def timed(name, fn):
        fn()
        ctx.sync()
        e0, e1 = C.c_void_p(), C.c_void_p()
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            fn()
        ctx.sync()
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        out[name] = round(ms.value / reps, 3)

    N, nang = 1024, 720
    g2 = lc.PolarGrid2D(synth.angles(nang), N)
    s2 = np.random.default_rng(0).standard_normal((8, nang, N)).astype(np.float32)
    timed("backproj2d_8x1024_ms", lambda: lc.backproj2d(s2, g2, (8, N, N)))
    dims, nproj, nrows = (32, 511, 511), 720, 128
    th = np.arange(nproj) * (2 * np.pi / nproj)
    g3 = lc.PolarGrid(th, np.full(nproj, -np.pi / 6), np.zeros(nproj), nrows, dims[2])
    p3 = np.random.default_rng(2).standard_normal(g3.dims()).astype(np.float32)
    timed("backproj3d_32x511_ms", lambda: lc.backproj(p3, g3, dims))
    out["note"] = "host-inclusive: each call copies its input in and its result out"
    print(json.dumps({"nufft_probe": out}), flush=True)


def pmc_probe(args):
    """Short run for rocprofv3 --pmc: a device copy of known size (calibration)
    and a few batched Toeplitz applies of the benchmark workload."""
    import ctypes as C
    import laminocam_amd as lc
    from laminocam_amd import synth
    hip = C.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    log(f"copy probe: {nbytes} bytes x3")
    ctx = lc.default_context()
    rng, Ns, As = workload(args.slices)
    geoms, rhs = build_problem(lc, synth, rng, Ns, As)
    batch = lc.Batch2D([geoms[(int(N), int(a))][1] for N, a in zip(Ns, As)], ctx=ctx)
    for i in range(len(Ns)):
        batch.set(i, batch.Y, rhs[i])
    prof = batch.profile_passes(args.kernel_reps)
    ctx.sync()
    for r in prof:
        log(f"pass N={r['N']} {PASS_NAMES[r['pass_']]} ms={r['ms']:.4f} bytes={r['bytes']:.0f}")
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=512)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--xtol", type=float, default=1e-5)
    ap.add_argument("--plain-slices", type=int, default=64,
                    help="slices of the workload solved one at a time for the plain-path rate")
    ap.add_argument("--no-err", action="store_true",
                    help="skip the per-class error check against the FP64 oracle")
    ap.add_argument("--cpu-iters", type=int, default=50)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--fp64-steps", type=int, default=2, help="timed FP64 batched solves")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 batched solve")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="served stream: requests/s")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="served stream: seconds of arrivals")
    ap.add_argument("--serve-sweep", type=str, default="50,100,200,400,800",
                    help="arrival rates (requests/s) of the mixed stream's rate sweep ('' skips)")
    ap.add_argument("--serve-2d-sweep", type=str, default="300,600,1200,2400",
                    help="arrival rates of the 2-D-only stream's rate sweep ('' skips)")
    ap.add_argument("--serve-sweep-seconds", type=float, default=2.0,
                    help="seconds of arrivals per sweep rate")
    ap.add_argument("--serve-3d-frac", type=float, default=0.05)
    ap.add_argument("--serve-2d-rate", type=float, default=600.0,
                    help="requests/s of the 2-D-only served stream (0: skip it)")
    ap.add_argument("--serve-lanes", type=int, default=4)
    ap.add_argument("--serve-budget-gb", type=float, default=96.0)
    ap.add_argument("--serve-max-batch", type=int, default=128,
                    help="served stream: most requests one lane batch takes")
    ap.add_argument("--sb-outer", type=int, default=20)
    ap.add_argument("--sb-steps", type=int, default=2,
                    help="timed batched Split-Bregman solves on the slices (0: skip)")
    ap.add_argument("--serve-plain-sample", type=int, default=60,
                    help="served stream: requests re-solved one at a time for the plain rate")
    ap.add_argument("--no-serve", action="store_true", help="skip the served stream")
    ap.add_argument("--no-lamino", action="store_true", help="skip the 3-D laminography solves")
    ap.add_argument("--lamino-iters", type=int, default=100)
    ap.add_argument("--no-lamino-cpu", action="store_true",
                    help="skip the CPU timing of the 3-D operator at the laminography shape")
    ap.add_argument("--lamino-only", action="store_true",
                    help="run only the 3-D laminography solves (one JSON line)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served stream and print its JSON (profiling)")
    ap.add_argument("--nufft-probe", action="store_true",
                    help="only the 2-D and 3-D back-projections (for counter runs of the spreads)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only run a calibration copy and a few Toeplitz applies (for rocprofv3 --pmc)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py: this build measures one GPU; --gpus {max(args.gpus, world)} is not "
              f"supported: the sharded multi-GPU path is not built", file=sys.stderr)
        sys.exit(2)

    if args.pmc_probe:
        pmc_probe(args)
        return
    if args.nufft_probe:
        nufft_probe()
        return

    import laminocam_amd as lc
    from laminocam_amd import synth

    if args.lamino_only:
        print(json.dumps({"lamino3d": lamino_leg(args, lc, synth, lc.default_context())}), flush=True)
        return
    def rates(txt):
        return [float(v) for v in txt.split(",") if v.strip()]

    if args.serve_only:
        served = served_stream(args, lc, synth, lc.default_context(), sweep=rates(args.serve_sweep))
        if args.serve_2d_rate > 0:
            served["two_d_only"] = served_stream(args, lc, synth, lc.default_context(),
                                                 frac3d=0.0, rate=args.serve_2d_rate,
                                                 sweep=rates(args.serve_2d_sweep))
        print(json.dumps({"served": served}), flush=True)
        return

    t_setup = time.perf_counter()
    ctx = lc.default_context()
    rng, Ns, As = workload(args.slices)
    geoms, rhs = build_problem(lc, synth, rng, Ns, As)
    psfs = [geoms[(int(N), int(a))][1] for N, a in zip(Ns, As)]
    batch = lc.Batch2D(psfs, ctx=ctx)
    for i in range(len(Ns)):
        batch.set(i, batch.Y, rhs[i])
    ctx.sync()
    log(f"setup {time.perf_counter() - t_setup:.1f} s: {len(Ns)} slices, {len(geoms)} geometries")

    # ---- timed batched solves ----
    for _ in range(args.warmup):
        batch.cg(args.iters, args.tol, args.xtol, graph=True)
    ctx.sync()
    step_s = []
    for _ in range(args.steps):
        ctx.sync()
        t0 = time.perf_counter()
        batch.cg(args.iters, args.tol, args.xtol, graph=True)
        ctx.sync()
        step_s.append(time.perf_counter() - t0)
    reps = batch.reports()
    iters = np.array([r["iters"] for r in reps])
    ms_step = 1e3 * float(np.mean(step_s))
    value = len(Ns) / float(np.mean(step_s))
    log(f"batched: {ms_step:.1f} ms/step, {value:.1f} slices/s, iterations mean {iters.mean():.1f}")
    xb = [batch.get(i, batch.X) for i in range(len(Ns))]

    # ---- dominant kernel: per-launch HIP-event timing on the library stream ----
    prof = batch.profile_passes(args.kernel_reps)
    # every pass (each Toeplitz class pass and the update) launches once per
    # CG iteration, so the dominant kernel is the one with the longest launch
    dom = max(prof, key=lambda r: r["ms"])
    dom_t = max((r for r in prof if r["pass_"] < 3), key=lambda r: r["ms"])
    tmap = {}
    tpath = os.path.join(ROOT, "profiles", "traffic_toeplitz_batched_cg.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tmap = json.load(f).get("bytes_per_launch", {})

    def kname(r):
        if r["pass_"] == 3:
            return (f"CG vector update k_cg_update (x += alpha p, r -= alpha Ap over all "
                    f"{r['nslices']} slices, 24 B per element)")
        return (f"toeplitz {PASS_NAMES[r['pass_']]} pass, N={r['N']} class "
                f"({r['nslices']} slices per launch)")

    def roof(r):
        a = r["bytes"] / (r["ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(a / HBM_PEAK_GBS, 4),
                "traffic": tmap.get(f"N{r['N']}_{PASS_NAMES[r['pass_']]}"),
                "kernel": kname(r), "launch_ms": round(r["ms"], 4), "algorithmic_bytes": r["bytes"]}
    roofline = roof(dom)
    roofline_toeplitz = roof(dom_t)
    achieved = roofline["achieved"]
    log(f"dominant: {roofline['kernel']}: {dom['ms']:.3f} ms/launch, {achieved:.0f} GB/s")

    # ---- FP64 path: the same batched solve in float64 ----
    fp64 = None
    if not args.no_fp64:
        b64 = lc.Batch2D(psfs, ctx=ctx, fp64=True)
        for i in range(len(Ns)):
            b64.set(i, b64.Y, rhs[i].astype(np.float64))
        b64.cg(args.iters, args.tol, args.xtol, graph=True)   # warm-up
        ctx.sync()
        t64 = []
        for _ in range(max(1, args.fp64_steps)):
            ctx.sync()
            t0 = time.perf_counter()
            b64.cg(args.iters, args.tol, args.xtol, graph=True)
            ctx.sync()
            t64.append(time.perf_counter() - t0)
        it64 = np.array([r["iters"] for r in b64.reports()])
        prof64 = b64.profile_passes(max(2, args.kernel_reps // 2))
        d64 = max(prof64, key=lambda r: r["ms"])
        fp64 = {"slices_per_s": round(len(Ns) / float(np.mean(t64)), 2),
                "ms_per_step": round(1e3 * float(np.mean(t64)), 3), "steps": len(t64),
                "cg_iterations_mean": round(float(it64.mean()), 2),
                "dominant_pass": f"{PASS_NAMES[d64['pass_']]} N={d64['N']}",
                "dominant_GBps": round(d64["bytes"] / (d64["ms"] * 1e-3) / 1e9, 1),
                "note": "FP64 vectors, PSF spectrum, FFTs and twiddles (cgsolver<double>)"}
        log(f"fp64 batched: {fp64['ms_per_step']} ms/step, {fp64['slices_per_s']} slices/s")
        b64.close()
        del b64

    # ---- plain path: one solve at a time through the C ABI, host arrays ----
    order = np.arange(len(Ns))
    plain_idx = order[:min(args.plain_slices, len(Ns))]
    lat = []
    max_diff = 0.0
    for i in plain_idx:
        t0 = time.perf_counter()
        x, _ = lc.cgsolver(psfs[i], rhs[i], np.zeros_like(rhs[i]), args.iters, args.tol, args.xtol)
        lat.append(time.perf_counter() - t0)
        max_diff = max(max_diff, float(np.max(np.abs(x - xb[i]))))
    lat = np.array(lat)
    # the plain rate for the whole workload: per-slice time by image size
    t_by_n = {}
    for i, t in zip(plain_idx, lat):
        t_by_n.setdefault(int(Ns[i]), []).append(t)
    t_mean = {n: float(np.mean(v)) for n, v in t_by_n.items()}
    if len(t_mean) == len(set(int(n) for n in Ns)):
        plain_total = sum(t_mean[int(n)] for n in Ns)
    else:
        plain_total = float(np.mean(lat)) * len(Ns)
    plain_rate = len(Ns) / plain_total
    log(f"plain: {plain_rate:.1f} slices/s over {len(plain_idx)} slices, max |x_plain - x_batched| "
        f"= {max_diff:.3e}")

    def rate_by_n(idx, times):
        tn = {}
        for i, t in zip(idx, times):
            tn.setdefault(int(Ns[i]), []).append(t)
        tn = {n: float(np.mean(v)) for n, v in tn.items()}
        if len(tn) == len(set(int(n) for n in Ns)):
            return len(Ns) / sum(tn[int(n)] for n in Ns)
        return len(Ns) / (float(np.mean(times)) * len(Ns))

    # the plain mode without its allocations and PCIe copies: each sampled
    # slice in a one-problem batch built before the clock, per-op launches with
    # the scalars read back after every iteration (LC_HOST_SYNC)
    lat_res, diff_res = [], 0.0
    for i in plain_idx:
        b1 = lc.Batch2D([psfs[i]], ctx=ctx)
        b1.set(0, b1.Y, rhs[i])
        ctx.sync()
        t0 = time.perf_counter()
        b1.cg(args.iters, args.tol, args.xtol, graph=False, host_sync=True)
        ctx.sync()
        lat_res.append(time.perf_counter() - t0)
        diff_res = max(diff_res, float(np.max(np.abs(b1.get(0, b1.X) - xb[i]))))
        b1.close()
    plain_res_rate = rate_by_n(plain_idx, lat_res)
    log(f"plain, device resident: {plain_res_rate:.1f} slices/s, max |diff| {diff_res:.3e}")

    # ---- Split-Bregman TV on the same slices: batched vs plain ----
    sb = None
    if args.sb_steps > 0:
        # zero-padded in-plane differences (the reference GPU path's boundary):
        # with the CPU path's interior-only differences a single slice has no
        # gradient and the shrink does nothing
        sb_kw = dict(lam=0.1, mu=10.0, outer=args.sb_outer, inner=1)
        batch.split_bregman(sb_kw["lam"], sb_kw["mu"], sb_kw["outer"], 1, args.tol, args.xtol,
                            graph=True, x0_zero=True, zero_pad=True)   # warm-up
        ctx.sync()
        t_sb = []
        for _ in range(args.sb_steps):
            ctx.sync()
            t0 = time.perf_counter()
            batch.split_bregman(sb_kw["lam"], sb_kw["mu"], sb_kw["outer"], 1, args.tol, args.xtol,
                                graph=True, x0_zero=True, zero_pad=True)
            ctx.sync()
            t_sb.append(time.perf_counter() - t0)
        outer_done = np.array([r["outer"] for r in batch.reports()])
        xsb = {int(i): batch.get(int(i), batch.X) for i in plain_idx}
        lat_sb, diff_sb = [], 0.0
        for i in plain_idx:
            t0 = time.perf_counter()
            x, _ = lc.split_bregman(psfs[i], rhs[i], np.zeros_like(rhs[i]), sb_kw["lam"],
                                    sb_kw["mu"], sb_kw["outer"], 1, args.tol, args.xtol,
                                    zero_pad=True)
            lat_sb.append(time.perf_counter() - t0)
            diff_sb = max(diff_sb, float(np.max(np.abs(x - xsb[int(i)]))))
        tn = {}
        for i, t in zip(plain_idx, lat_sb):
            tn.setdefault(int(Ns[i]), []).append(t)
        tn = {n: float(np.mean(v)) for n, v in tn.items()}
        if len(tn) == len(set(int(n) for n in Ns)):
            sb_plain_total = sum(tn[int(n)] for n in Ns)
        else:
            sb_plain_total = float(np.mean(lat_sb)) * len(Ns)
        sb_rate = len(Ns) / float(np.mean(t_sb))
        sb_plain_rate = len(Ns) / sb_plain_total
        sb = {"slices_per_s": round(sb_rate, 2), "ms_per_step": round(1e3 * float(np.mean(t_sb)), 3),
              "plain_slices_per_s": round(sb_plain_rate, 2),
              "speedup_vs_plain": round(sb_rate / sb_plain_rate, 2),
              "outer_mean": round(float(outer_done.mean()), 2),
              "max_abs_diff_plain_vs_batched": diff_sb,
              "boundary": "zero_pad (LC_TV_ZERO_PAD: the reference GPU path's zero-padded "
                          "in-plane differences, under which the shrink acts on single slices)",
              "workload": f"the {len(Ns)} slices above, Split-Bregman isotropic TV, lambda "
                          f"{sb_kw['lam']}, mu {sb_kw['mu']}, {sb_kw['outer']} outer x 1 inner, "
                          f"x0 = 0; plain: the first {len(plain_idx)} slices one at a time"}
        log(f"split-bregman: batched {sb_rate:.1f} slices/s, plain {sb_plain_rate:.1f} slices/s, "
            f"max |diff| {diff_sb:.3e}")

    # ---- relative error vs the FP64 oracle, one slice per size class ----
    err = None
    if not args.no_err:
        err = rel_err_by_class(args, lc, synth, Ns, As, psfs, rhs, xb)
        sm = err["summary"]
        log(f"rel err vs FP64 oracle, {len(err['classes'])} classes: within 1e-3 up to "
            f"{sm['fp32_horizon_1e-3_min_max']} iterations (FP32), {sm['fp64_horizon_1e-3_min_max']} "
            f"(FP64); at {args.iters} it FP32 within 1e-3 in {sm['fp32_within_1e-3_at_full_horizon']}, "
            f"FP64 in {sm['fp64_gpu_within_1e-3_at_full_horizon']}; FP32 / reference float path "
            f"error ratio (median) {sm['fp32_over_reference_float_path_median']}")

    # ---- served mixed stream ----
    served = None
    if not args.no_serve:
        served = served_stream(args, lc, synth, ctx, sweep=rates(args.serve_sweep))
        log(f"served: {served['requests']} requests, {served['slices_per_s']} slices/s, latency "
            f"p50 {served['latency_ms']['p50']} ms p99 {served['latency_ms']['p99']} ms, "
            f"{served['batches']} batches")
        if args.serve_2d_rate > 0:
            # the same stream without 3-D volumes: the 2-D slices alone
            served["two_d_only"] = served_stream(args, lc, synth, ctx, frac3d=0.0,
                                                 rate=args.serve_2d_rate,
                                                 sweep=rates(args.serve_2d_sweep))
            s2 = served["two_d_only"]
            log(f"served 2-D only: {s2['slices_per_s']} slices/s (capacity "
                f"{s2['capacity']['slices_per_s']}), plain {s2['plain_same_mix']['slices_per_s']}, "
                f"p50 {s2['latency_ms']['p50']} ms p99 {s2['latency_ms']['p99']} ms")

    # ---- 3-D laminography solves ----
    lamino = None
    if not args.no_lamino:
        lamino = lamino_leg(args, lc, synth, ctx)
        log(f"lamino 3-D: CG {lamino['cg']['iters']} it {lamino['cg']['seconds']} s, SB "
            f"{lamino['split_bregman']['outer']} outer {lamino['split_bregman']['seconds']} s, "
            f"PSF {lamino['psf_build_s']} s")

    # ---- CPU baseline ----
    cpu = None
    if not args.no_cpu:
        rhs_by_n = {}
        for i in range(len(Ns)):
            N = int(Ns[i])
            if N not in rhs_by_n or As[i] < rhs_by_n[N][2]:
                rhs_by_n[N] = (rhs[i], synth.angles(int(As[i])), int(As[i]))
        rhs_by_n = {k: v[:2] for k, v in rhs_by_n.items()}
        cpu = cpu_baseline(args, iters, Ns, rhs_by_n)
        log(f"cpu baseline: {cpu['value']:.3f} slices/s on {cpu['cores']} threads")

    out = {
        "metric": "slices/sec + p50/p99 latency, batched/served vs plain GPU CG at <=1e-3 rel err",
        "value": round(value, 2),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 dots/norms/PSF/twiddles)",
        "data": "synthetic (seeded random-ellipse phantoms, NUFFT sinograms, 1% Gaussian noise)",
        "config": {
            "workload": f"batched 2-D CG: {len(Ns)} slices, N in {list(N_SET)}, angles in "
                        f"{list(ANG_SET)} over [0,180), nrad = N, {args.iters} iterations, "
                        f"tol {args.tol:g}, xtol {args.xtol:g}, one hipGraph replay per iteration",
            "slices": len(Ns), "iters": args.iters,
        },
        "latency_ms": {"p50": round(1e3 * float(np.median(step_s)), 3),
                       "p99": round(1e3 * float(np.percentile(step_s, 99)), 3),
                       "note": "batched: every slice is submitted at the start of a step and "
                               "completes with it"},
        "plain": {"slices_per_s": round(plain_rate, 2), "sample_slices": int(len(plain_idx)),
                  "latency_ms": {"p50": round(1e3 * float(np.median(lat)), 3),
                                 "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
                  "max_abs_diff_vs_batched": max_diff,
                  "device_resident": {
                      "slices_per_s": round(plain_res_rate, 2),
                      "max_abs_diff_vs_batched": diff_res,
                      "note": "the same slices in one-problem batches built before the clock: "
                              "per-op launches, scalars read back every iteration, no "
                              "allocation or PCIe copies in the timed region"}},
        "speedup_vs_plain": round(value / plain_rate, 2),
        "speedup_vs_plain_device_resident": round(value / plain_res_rate, 2),
        "cg_iterations": {"mean": round(float(iters.mean()), 2), "min": int(iters.min()),
                          "max": int(iters.max())},
        "rel_err_vs_fp64": err,
        "config1": err["config1"] if err else None,
        "fp64_batched": fp64,
        "split_bregman_batched": sb,
        "served": served,
        "lamino3d": lamino,
        "roofline": roofline,
        "roofline_toeplitz": roofline_toeplitz,
        "hbm_copy_reference": copy_reference(),
        "cpu_baseline": cpu,
        "kernel_passes": [{"N": r["N"], "pass": PASS_NAMES[r["pass_"]], "ms": round(r["ms"], 4),
                           "GBps": round(r["bytes"] / (r["ms"] * 1e-3) / 1e9, 1)} for r in prof],
    }
    print(json.dumps(out), flush=True)


def copy_reference(nbytes=1 << 30, reps=10):
    """Device-to-device hipMemcpyAsync of nbytes on the null stream, timed with
    HIP events: the HBM rate the runtime's own copy reaches on this GPU, a
    practical ceiling for kernels that read and write equal amounts."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        return None
    try:
        hip.hipMemset(a, 0, C.c_size_t(nbytes))
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        d2d = 3   # hipMemcpyDeviceToDevice
        for _ in range(3):
            hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), d2d, None)
        hip.hipDeviceSynchronize()
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), d2d, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        gbps = 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
        return {"GB/s": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBS, 4),
                "how": f"hipMemcpyAsync device to device, {nbytes >> 20} MiB, {reps} copies, "
                       "HIP events (read + write bytes)"}
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(a)
        hip.hipFree(b)


if __name__ == "__main__":
    main()
