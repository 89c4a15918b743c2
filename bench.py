#!/usr/bin/env python3
"""Reconstructions per second and p50/p99 latency of depth-8 screened Poisson
reconstruction on one MI355X, FP32-mixed against FP64.

Workload: 200,000 oriented points on a torus (R = 0.6, r = 0.25), Gaussian position noise
0.002 x bbox diagonal, 5 degree normal noise, normal lengths U(0.5, 1.5) with confidence
weighting; depth 8 (257^3 grid), reference defaults (point weight 4, 4 V-cycles of 3/3
smoothing, 40 coarsest iterations, 1.5 samples per node, scale 1.1).

Headline (`value`): a "step" is one batched graph launch of --batch such reconstructions
(different seeds) in mixed precision; --warmup launches, then exactly --steps timed launches.
Points are resident in HBM before the timed region; each launch filters and normalises
them, splats, assembles, solves, computes the iso value and extracts the mesh on the device
and returns the counts (the meshes stay in HBM). `value` = batch x steps / seconds; p50/p99
are the launch latencies (every request of a batch completes with it).

Also reported:
  modes        device-resident single requests: FP64 (the plain GPU solve), mixed, mixed graph
  end_to_end   NumPy points in, Mesh out: reconstruct() FP64 and mixed (the drop-in call, host
               transfers included: meshes come back through pinned host buffers), and the
               batched form with its host uploads and downloads
  accuracy     mixed vs FP64 mesh distance (exact point-to-triangle, GPU) over the diagonal
  roofline     the dominant kernel: the fused smoothing of the batch's finest level (two
               red-black iterations of all instances in one z-marching pass)
  batch64      64 clouds of uneven depth and size, mixed and FP64, per-instance error vs FP64
  depth9       2M-point superquadric at depth 9, mixed and FP64
  serve        a QPS sweep of the served stream, then --serve-requests requests at a rate below
               the measured capacity, with a sampled per-result error against FP64
  cpu_baseline the CPU oracle (the reference algorithm in C) on the host: 1 thread and all
               threads, best of 3 and p50, on the 2k-point depth-5 sphere and this workload
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mojo-poissonrecon_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the roofline kernel (kb_march<4, 33>, two red-black iterations of
# the 257^3 finest level of all instances of the batch), keyed by (depth, batch):
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from rocprofv3 counter runs (the timing repetitions'
# launches), the factors calibrated on a plain elementwise copy
# (profiles/kb_march_f32_depth8_pmc.txt).
ROOFLINE_TRAFFIC_BYTES = {(8, 8): 2_228_616_096}


def log(*args):
    print("[bench]", *args, file=sys.stderr, flush=True)


def timed(step, steps: int, warmup: int):
    for _ in range(warmup):
        step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        step()
        lat.append(time.perf_counter() - s)
    return time.perf_counter() - t0, np.array(lat)


def pct(lat, q):
    return round(float(np.percentile(lat, q)) * 1e3, 4) if len(lat) else None


def diag_of(p):
    return float(np.linalg.norm(p.max(axis=0) - p.min(axis=0)))


def mesh_error(a, b, diag) -> dict:
    """exact vertex-to-surface distances between two meshes, over the bbox diagonal"""
    from mojo_poissonrecon.metrics import mesh_distance

    d = mesh_distance(a, b)
    return {"hausdorff": d.hausdorff / diag, "chamfer": d.chamfer / diag}


# ---- CPU baseline -----------------------------------------------------------------------

def cpu_leg(points, normals, depth, confidence, handle, budget_s, runs) -> dict:
    import oracle

    times = []
    start = time.perf_counter()
    while len(times) < runs and (len(times) < 3 or time.perf_counter() - start < budget_s):
        t = time.perf_counter()
        oracle.reconstruct(points, normals, depth=depth, confidence=confidence, handle=handle)
        times.append(time.perf_counter() - t)
    t = np.array(times)
    return {"recon_per_s_best_of_3": round(1.0 / t[:3].min(), 4),
            "recon_per_s_p50": round(1.0 / float(np.median(t)), 4),
            "runs": len(t), "p50_ms": round(float(np.median(t)) * 1e3, 2)}


def cpu_baseline(points, normals, depth: int) -> tuple[dict, dict]:
    """The oracle on the host: the serial build (the reference's parallelize is a serial
    loop) and the OpenMP build on every thread, on the 2k-point depth-5 sphere and on the
    headline workload. Returns (cpu_baseline object, all legs)."""
    import oracle
    from mojo_poissonrecon import shapes

    omp = oracle.openmp()
    threads = int(omp.oracle_threads())
    sp, sn = shapes.sphere(2000, seed=0)
    legs = {
        "host_cores": len(os.sched_getaffinity(0)),
        "openmp_threads": threads,
        "sphere_2k_depth5": {"1_thread": cpu_leg(sp, sn, 5, False, None, 5.0, 10),
                             f"{threads}_threads": cpu_leg(sp, sn, 5, False, omp, 5.0, 10)},
        f"torus_{len(points)}_depth{depth}": {
            "1_thread": cpu_leg(points, normals, depth, True, None, 28.0, 10),
            f"{threads}_threads": cpu_leg(points, normals, depth, True, omp, 12.0, 10)},
        "published_reference": "24.075 ms for 2k sphere points at depth 5 (41.5 recon/s), "
                               "Xeon E5-2697 v4, README.md:121 - a different machine",
    }
    one = legs[f"torus_{len(points)}_depth{depth}"]["1_thread"]
    base = {"value": one["recon_per_s_best_of_3"], "unit": "recon/s", "cores": 1,
            "kind": "port",
            "sample": f"full FP64 depth-{depth} reconstructions of the same {len(points)}-point "
                      f"torus by the serial C oracle, best of 3 of {one['runs']} runs "
                      f"(p50 {one['recon_per_s_p50']} recon/s); OpenMP and 2k-sphere legs in "
                      "cpu_baseline_legs"}
    return base, legs


# ---- device modes -----------------------------------------------------------------------

def make_batch(clouds, depths, precision, graph=True, host_mesh=False, inner_cycles=0):
    """Jobs and a batch over the clouds. host_mesh: the extraction also copies each mesh to
    the job's pinned host buffers (the end-to-end legs); without, meshes stay in HBM.
    inner_cycles: FP32 V-cycles per FP64 defect (mixed; 0 the library's default, one)."""
    from mojo_poissonrecon.batch import Batch
    from mojo_poissonrecon.core import Job, prepare_reconstruct

    reqs = [prepare_reconstruct(p, n, depth=d, precision=precision, confidence=True,
                                inner_cycles=inner_cycles)
            for (p, n), d in zip(clouds, depths)]
    jobs = [Job(r.depth, len(r.points), r.precision, host_mesh=host_mesh) for r in reqs]
    for j, r in zip(jobs, reqs):
        j.set_points(r.points, r.normals)
    b = Batch(jobs)
    if graph:
        b.capture([r.params for r in reqs])
    return b, jobs, reqs


def run_batch_launches(b, steps, warmup):
    def step():
        b.launch()
        b.wait()

    total, lat = timed(step, steps, warmup)
    return total / steps, lat


def batch_kernels(b, reps, kernels):
    from mojo_poissonrecon import _lib

    out = {}
    for name, kid in kernels:
        ms, nbytes = _lib.ctypes.c_double(), _lib.ctypes.c_double()
        _lib.check(_lib.lib().mpr_batch_time_kernel(b.handle, kid, reps, _lib.ctypes.byref(ms),
                                                     _lib.ctypes.byref(nbytes)),
                   "mpr_batch_time_kernel")
        out[name] = dict(ms=ms.value, bytes=nbytes.value,
                         gbs=nbytes.value / (ms.value * 1e-3) / 1e9)
    return out


def time_kernels(jobs, reps: int) -> dict:
    """Dominant kernels at the finest level, timed with HIP events on each job's own stream
    (the stream the pipeline launches them on)."""
    from mojo_poissonrecon import _lib

    kernels = {}
    for kname, kid, job in (("rb_sweep_f64", 8, jobs["fp64"]), ("rb_sweep_f32", 7, jobs["mixed"]),
                            ("residual_f32", 9, jobs["mixed"]), ("residual_f64", 2, jobs["fp64"]),
                            ("restrict_f64", 3, jobs["fp64"]), ("prolong_f64", 4, jobs["fp64"]),
                            ("defect_f64_f32", 6, jobs["mixed"]),
                            ("restrict_f32", 10, jobs["mixed"]), ("prolong_f32_c1", 11, jobs["mixed"]),
                            ("restrict_f64_batched", 12, jobs["fp64"]),
                            ("defect_f64_f32_batched", 13, jobs["mixed"])):
        ms, nbytes = _lib.ctypes.c_double(), _lib.ctypes.c_double()
        _lib.check(_lib.lib().mpr_job_time_kernel(job.handle, kid, reps, _lib.ctypes.byref(ms),
                                                   _lib.ctypes.byref(nbytes)),
                   "mpr_job_time_kernel")
        kernels[kname] = dict(ms=ms.value, bytes=nbytes.value,
                              gbs=nbytes.value / (ms.value * 1e-3) / 1e9)
    return kernels


def end_to_end(points, normals, clouds, depth, steps, warmup) -> dict:
    """NumPy in, Mesh out: the drop-in reconstruct() (validation, upload, device pass, mesh
    download) per request, and the batched form with persistent jobs (upload every cloud,
    one launch, download every mesh) per batch."""
    import mojo_poissonrecon as mpr
    from mojo_poissonrecon.batch import meshes_from_jobs, upload_all

    out = {}
    for prec in ("fp64", "mixed"):
        total, lat = timed(lambda: mpr.reconstruct(points, normals, depth=depth, confidence=True,
                                                   precision=prec), steps, warmup)
        out[f"reconstruct_{prec}"] = {"recon_per_s": round(steps / total, 3), "p50_ms": pct(lat, 50),
                                      "p99_ms": pct(lat, 99)}
    b, jobs, reqs = make_batch(clouds, [depth] * len(clouds), "mixed", graph=False,
                               host_mesh=True)

    def step():
        upload_all(jobs, clouds)
        b.run([r.params for r in reqs])
        return meshes_from_jobs(jobs, b.wait())

    total, lat = timed(step, steps, warmup)
    out["batch_mixed"] = {"recon_per_s": round(len(clouds) * steps / total, 3),
                          "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "batch": len(clouds)}
    b.close()
    for j in jobs:
        j.close()
    mpr.release_device_memory()
    return out


def batch64(instances: int) -> dict:
    from mojo_poissonrecon import shapes
    from mojo_poissonrecon.core import mesh_from_job

    clouds, depths = shapes.uneven_batch_workload(instances, seed=1)
    rates, meshes = {}, {}
    launch_ms = None
    for prec in ("mixed", "fp64"):
        b, jobs, reqs = make_batch(clouds, depths, prec)
        per, _ = run_batch_launches(b, 3, 1)
        rates[prec] = round(instances / per, 2)
        meshes[prec] = [mesh_from_job(j, r) for j, r in zip(jobs, b.wait())]
        if prec == "mixed":
            launch_ms = round(per * 1e3, 3)
        b.close()
        for j in jobs:
            j.close()
    errs = np.array([mesh_error(a, c, diag_of(p))["hausdorff"]
                     for (p, _), a, c in zip(clouds, meshes["mixed"], meshes["fp64"])])
    return {"instances": instances, "depths": "U{5,6,7,8}", "points": "2k-200k log-uniform",
            "mixed_recon_per_s": rates["mixed"], "fp64_recon_per_s": rates["fp64"],
            "launch_ms": launch_ms,
            "hausdorff_mixed_vs_fp64_over_diag": {"max": float(errs.max()),
                                                  "p50": float(np.median(errs)),
                                                  "p99": float(np.percentile(errs, 99))}}


def depth9_run(steps: int) -> dict:
    """2,000,000 superquadric points (exponents 0.3/0.3, half-extents 0.8/0.6/0.4, sigma
    0.001 diag, 3 deg normal noise, unit normals), mixed and all-FP64 single requests at
    depth 9, with both final FP64 residual norms."""
    from mojo_poissonrecon import shapes
    from mojo_poissonrecon.core import Job, mesh_from_job, prepare_reconstruct

    p, n = shapes.superquadric_workload(2_000_000, seed=0)
    out, meshes = {}, {}
    for name in ("fp64", "mixed"):
        req = prepare_reconstruct(p, n, depth=9, precision=name)
        job = Job(9, len(p), req.precision, host_mesh=False)
        job.set_points(req.points, req.normals)

        def step(job=job, req=req):
            job.run(req.params)
            job.wait()

        total, lat = timed(step, steps, 1)
        r = job.wait()
        meshes[name] = mesh_from_job(job, r)
        out[name] = {"recon_per_s": round(steps / total, 3), "p50_ms": pct(lat, 50),
                     "residual_norm": r.residual_norm, "iso_value": r.iso_value,
                     "vertices": int(r.vertex_count),
                     "footprint_gb": round(Job.footprint(9, len(p), req.precision) / 1e9, 2)}
        job.close()
    err = mesh_error(meshes["mixed"], meshes["fp64"], diag_of(p))
    return {"workload": "superquadric 2,000,000 points, depth 9 (513^3), sigma 0.001 diag, "
                        "3 deg normal noise", **out,
            "hausdorff_mixed_vs_fp64_over_diag": err["hausdorff"],
            "chamfer_mixed_vs_fp64_over_diag": err["chamfer"],
            "residual_ratio_mixed_over_fp64": out["mixed"]["residual_norm"]
            / out["fp64"]["residual_norm"]}


def new_server():
    from mojo_poissonrecon.serve import ReconstructionServer

    return ReconstructionServer(max_inflight=32, max_batch=8, precision="mixed")


def serve_stream(requests: int, qps: float, seed: int, warmup: int = 0,
                 sample_every: int = 0, server=None) -> dict:
    """The served stream (shapes.served_schedule: Poisson arrivals at `qps`, 5k-500k points,
    three shape families randomly rotated and scaled, depths 6:7:8:9 = 4:3:2:1), mixed
    precision. The first `warmup` requests fill the server's job pool and are excluded from
    the latency percentiles. With sample_every, every such request's mesh is compared with
    its all-FP64 solo solve. `server`: a running server to use (its job pool stays warm
    from earlier streams; counts below are this stream's), else a new one, closed after."""
    import mojo_poissonrecon as mpr
    from mojo_poissonrecon import shapes

    from collections import deque
    from concurrent.futures import ThreadPoolExecutor

    total = requests + warmup
    base, items = shapes.served_schedule(total, qps, seed=seed)
    own = server is None
    if own:
        server = new_server()
    st0 = dict(vars(server.stats))
    st0.pop("latencies")
    done, clouds = [], {}
    id0 = None
    # the client side: 4 threads form each request's rows (rotation and scaling of a pool
    # cloud) up to 32 requests ahead of their arrival, off the server's thread; a request
    # whose rows are late is submitted when they are ready, its latency still counted from
    # its scheduled arrival
    client = ThreadPoolExecutor(4)
    ahead: deque = deque()
    formed = 0

    def refill():
        nonlocal formed
        while formed < total and len(ahead) < 32:
            ahead.append(client.submit(shapes.served_cloud, base, items[formed]))
            formed += 1

    refill()
    start = time.perf_counter()
    submitted = 0
    last_log = start
    while len(done) < total:
        now = time.perf_counter() - start
        while ahead and items[submitted][0] <= now and ahead[0].done():
            arrival, p, n, d = ahead.popleft().result()
            r = server.submit(p, n, depth=d, arrival=start + arrival)
            id0 = r.id if id0 is None else id0
            if sample_every and submitted >= warmup and (submitted - warmup) % sample_every == 0:
                clouds[r.id] = (p, n, d)
            submitted += 1
            refill()
        done += server.step()
        if time.perf_counter() - last_log > 20.0:
            last_log = time.perf_counter()
            log(f"serve {qps:g} QPS: {len(done)}/{total} done")
    wall = time.perf_counter() - start
    client.shutdown()
    stats = server.stats
    delta = {k: getattr(stats, k) - v for k, v in st0.items()
             if not k.startswith("peak")}
    if own:
        server.close()
    measured = [r for r in done if r.id >= id0 + warmup]
    lat = np.array([r.latency for r in measured])
    last = max(r.completed for r in measured) - start
    first = min(r.arrival for r in measured) - start
    out = {"requests": requests, "warmup_requests": warmup, "qps_offered": qps,
           "recon_per_s": round(len(measured) / (last - first), 2),
           "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "batches": delta["batches"],
           "batches_reused": delta["batches_reused"], "failed": delta["failed"],
           "fallbacks": delta["fallbacks"], "peak_inflight": stats.peak_inflight,
           "peak_hbm_gb": round(stats.peak_bytes / 1e9, 2), "jobs_created": delta["jobs_created"],
           "host_s": {"prepare": round(delta["host_prepare"], 3),
                      "launch": round(delta["host_launch"], 3),
                      "collect": round(delta["host_collect"], 3), "wall": round(wall, 3)}}
    if clouds:
        errs = []
        for r in measured:
            if r.id in clouds:
                p, n, d = clouds[r.id]
                ref = mpr.reconstruct(p, n, depth=d, precision="fp64")
                errs.append(mesh_error(r.mesh, ref, diag_of(p))["hausdorff"])
        mpr.release_device_memory()
        out["sampled_hausdorff_mixed_vs_fp64_over_diag"] = {
            "results": len(errs), "max": float(np.max(errs)), "p50": float(np.median(errs))}
    return out


def serve_runs(requests: int, sweep: list[float], sweep_requests: int, qps: float | None):
    """A QPS sweep of short streams (capacity = the best completed rate), then the long
    stream at `qps`, or at 80 % of that capacity, all on one server: its job pool warms up
    in the first stream and stays warm, as in a long-running server (a new server's first
    few hundred requests create most of its jobs, each a device allocation of up to tens
    of GB, and measure that warm-up instead)."""
    server = new_server()
    points = []
    for q in sweep:
        r = serve_stream(sweep_requests, q, seed=5, warmup=50, server=server)
        points.append({k: r[k] for k in ("qps_offered", "recon_per_s", "p50_ms", "p99_ms")})
        log("serve sweep:", points[-1])
    capacity = max(p["recon_per_s"] for p in points) if points else None
    rate = qps or (round(0.8 * capacity, 1) if capacity else 100.0)
    main = serve_stream(requests, rate, seed=2, warmup=100, sample_every=max(1, requests // 20),
                        server=server)
    server.close()
    log("serve:", main)
    return {"sweep": points, "capacity_recon_per_s": capacity, "stream": main,
            "depth_mix": "6:7:8:9 = 4:3:2:1", "points": "5k-500k log-uniform"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batched launches")
    ap.add_argument("--warmup", type=int, default=3, help="untimed batched launches first")
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--points", type=int, default=200_000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=10_000)
    ap.add_argument("--serve-qps", type=float, default=None,
                    help="offered rate of the long served stream (default: 80%% of the "
                         "capacity the sweep measures)")
    ap.add_argument("--serve-sweep", default="100,200,400,800",
                    help="comma-separated offered rates of the short sweep streams ('' for none)")
    ap.add_argument("--serve-sweep-requests", type=int, default=1000)
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served streams and print their line")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 64-batch, depth-9 and serving runs")
    ap.add_argument("--kernel-reps", type=int, default=40)
    ap.add_argument("--no-autotune", action="store_true",
                    help="launch the batched kernels with the default workgroup size")
    ap.add_argument("--only", choices=("fp64", "mixed", "mixed_batch"),
                    help="time one mode and print its line (for per-mode kernel traces)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="one reconstruction per precision, then only the kernel timings "
                         "(a short program for hardware-counter runs)")
    ap.add_argument("--march", type=int, choices=(0, 1, 2), default=None,
                    help="red-black iterations per fused smoothing pass of the 257^3 and larger "
                         "FP32 levels (0: single colour sweeps; default: the library's, 2); "
                         "0 only with --only, since the roofline kernel is the fused pass")
    args = ap.parse_args()
    if args.march == 0 and not args.only:
        print("--march 0 leaves no fused pass to report as the roofline kernel; use it with "
              "--only", file=sys.stderr)
        return 2
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent reconstructions shard one per GPU with no exchange)", file=sys.stderr)
        return 2

    import mojo_poissonrecon as mpr
    from mojo_poissonrecon import _build, _lib, shapes
    from mojo_poissonrecon.core import Job, mesh_from_job, prepare_reconstruct

    if not os.path.exists(_build.LIB):
        mpr.build()
    if args.march is not None:
        from mojo_poissonrecon import tune as march_tune
        march_tune.set_march_iterations(args.march)
    if not args.no_autotune:
        from mojo_poissonrecon import tune
        t0 = time.perf_counter()
        chosen = tune.autotune(max_depth=max(5, min(args.depth, 10)), reps=20)
        log(f"autotune {time.perf_counter() - t0:.2f} s:",
            " ".join(f"{k}/{p}/{s}={b}x{g}x{z}" for (k, p, s), (b, g, z) in sorted(chosen.items())))

    sweep = [float(x) for x in args.serve_sweep.split(",") if x.strip()]
    if args.serve_only:
        print(json.dumps({"mode": "serve", **serve_runs(args.serve_requests, sweep,
                                                        args.serve_sweep_requests,
                                                        args.serve_qps)}), flush=True)
        return 0

    points, normals = shapes.torus_workload(args.points, seed=0)
    bclouds = [shapes.torus_workload(args.points, seed=s) for s in range(args.batch)]
    sweep_kernels = [("march_f32", 14), ("rb_sweep_f32", 7), ("residual_f32", 9),
                     ("restrict_f32", 10), ("prolong_f32_c1", 11), ("defect_f64_f32", 13)]
    if args.march == 0:
        sweep_kernels = sweep_kernels[1:]
    if args.only == "mixed_batch":
        b, jobs, _ = make_batch(bclouds, [args.depth] * args.batch, "mixed")
        per, blat = run_batch_launches(b, args.steps, args.warmup)
        bk = batch_kernels(b, args.kernel_reps, sweep_kernels)
        print(json.dumps({"mode": "mixed_batch", "batch": args.batch,
                          "recon_per_s": round(args.batch / per, 3),
                          "p50_ms": pct(blat, 50), "steps": args.steps, "warmup": args.warmup,
                          "kernels_batch": {k: {"us": round(v["ms"] * 1e3, 2),
                                                "GBps": round(v["gbs"], 1),
                                                "bytes": round(v["bytes"])}
                                            for k, v in bk.items()}}), flush=True)
        return 0

    # device-resident single requests (points uploaded once, before the timed region)
    modes, meshes, jobs = {}, {}, {}
    runs = (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED))
    if args.only:
        runs = tuple(r for r in runs if r[0] == args.only)
    req = prepare_reconstruct(points, normals, depth=args.depth, confidence=True)
    for name, prec in runs:
        job = Job(args.depth, len(points), prec, host_mesh=False)
        job.set_points(req.points, req.normals)

        def step(job=job):
            job.run(req.params)
            job.wait()

        total, lat = timed(step, args.steps, args.warmup)
        r = job.wait()
        meshes[name] = mesh_from_job(job, r)
        modes[name] = dict(value=args.steps / total, p50=pct(lat, 50), p99=pct(lat, 99),
                           nv=int(r.vertex_count))
        jobs[name] = job
        log(f"{name}: {modes[name]['value']:.2f} recon/s, p50 {modes[name]['p50']} ms")
    if args.only:
        m = modes[args.only]
        print(json.dumps({"mode": args.only, "recon_per_s": round(m["value"], 3),
                          "p50_ms": m["p50"], "p99_ms": m["p99"], "vertices": m["nv"],
                          "steps": args.steps, "warmup": args.warmup}), flush=True)
        for job in jobs.values():
            job.close()
        return 0
    if args.kernels_only:
        kernels = time_kernels(jobs, args.kernel_reps)
        print(json.dumps({k: {"us": round(v["ms"] * 1e3, 2), "GBps": round(v["gbs"], 1)}
                          for k, v in kernels.items()}), flush=True)
        for job in jobs.values():
            job.close()
        return 0
    jm = jobs["mixed"]
    jm.capture(req.params)

    def gstep():
        jm.launch()
        jm.wait()

    total, lat = timed(gstep, args.steps, args.warmup)
    modes["mixed_graph"] = dict(value=args.steps / total, p50=pct(lat, 50), p99=pct(lat, 99))
    log(f"mixed_graph: {modes['mixed_graph']['value']:.2f} recon/s")
    accuracy = mesh_error(meshes["mixed"], meshes["fp64"], diag_of(points))
    kernels = time_kernels(jobs, args.kernel_reps)
    for job in jobs.values():
        job.close()

    # headline: --batch depth-8 reconstructions per graph launch, exactly --steps launches
    b, bjobs, _ = make_batch(bclouds, [args.depth] * args.batch, "mixed")
    per_launch, blat = run_batch_launches(b, args.steps, args.warmup)
    bkernels = batch_kernels(b, args.kernel_reps, sweep_kernels)
    b.close()
    for j in bjobs:
        j.close()
    modes["mixed_batch"] = dict(value=args.batch / per_launch, p50=pct(blat, 50),
                                p99=pct(blat, 99))
    log(f"mixed_batch x{args.batch}: {modes['mixed_batch']['value']:.2f} recon/s")
    # the same batch with two FP32 V-cycles per FP64 defect (inner_cycles = 2): two FP64
    # defects and corrections fewer, and the first cloud's mesh error against its FP64 mesh
    b, bjobs, _ = make_batch(bclouds, [args.depth] * args.batch, "mixed", inner_cycles=2)
    per2, blat2 = run_batch_launches(b, args.steps, args.warmup)
    mesh2 = mesh_from_job(bjobs[0], b.wait()[0])
    b.close()
    for j in bjobs:
        j.close()
    inner2 = {"value": round(args.batch / per2, 3), "p50_ms": pct(blat2, 50),
              "p99_ms": pct(blat2, 99),
              "accuracy_vs_fp64_over_diag": mesh_error(mesh2, meshes["fp64"], diag_of(points))}
    log(f"mixed_batch x{args.batch}, inner_cycles 2: {inner2['value']:.2f} recon/s")
    e2e = end_to_end(points, normals, bclouds, args.depth, max(5, args.steps // 2), 2)
    log("end_to_end:", e2e)

    # the dominant kernel: the fused smoothing of all instances' finest levels, two red-black
    # iterations per launch. Algorithmic bytes by the rule for a red-black iteration (4 words
    # per node: x read, rhs and screen read, x written, whether or not iterations are fused):
    # 8 n^3 words per instance. The pass itself loads each plane's input colour, rhs and live
    # screening words once and writes both colours once (loaded_bytes), so its achieved rate
    # by the rule can exceed what a single sweep reaches; traffic is the counters' bytes.
    dom = bkernels["march_f32"]
    n = (1 << args.depth) + 1
    alg_bytes = float(args.batch) * 8.0 * n ** 3 * 4.0
    achieved = alg_bytes / (dom["ms"] * 1e-3) / 1e9
    traffic = ROOFLINE_TRAFFIC_BYTES.get((args.depth, args.batch))

    extras = {}
    if not args.no_extras:
        extras["batch64"] = batch64(args.instances)
        log("batch64:", extras["batch64"])
        extras["depth9"] = depth9_run(3)
        log("depth9:", extras["depth9"])
        extras["serve"] = serve_runs(args.serve_requests, sweep, args.serve_sweep_requests,
                                     args.serve_qps)

    mixed_b, fp64 = modes["mixed_batch"], modes["fp64"]
    line = {
        "metric": "recon/s + p50/p99 latency, depth 8, FP32-mixed vs FP64 at fixed Hausdorff",
        "value": round(mixed_b["value"], 3),
        "unit": "recon/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_launch * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycles in f64 defect correction (f64 splat, residual, iso)",
        "inner_cycles": "library default: one f32 V-cycle per f64 defect (4 outer steps); "
                        "mixed_batch_inner_cycles_2 is the same batch with two",
        "data": "synthetic",
        "config": {
            "workload": f"torus {len(points)} points, depth {args.depth} ({n}^3), sigma 0.002 "
                        "diag, 5 deg normal noise, normal lengths U(0.5,1.5), confidence on",
            "mode": f"mixed precision, {args.batch} reconstructions per batched graph launch "
                    "(a step); points and meshes resident in HBM",
            "depth": args.depth,
            "points": len(points),
            "parallelism": "single GPU",
        },
        "p50_ms": mixed_b["p50"],
        "p99_ms": mixed_b["p99"],
        "latency_note": "p50/p99 of a batched launch: every request of the batch completes "
                        "with it",
        "modes": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                      for kk, vv in v.items() if kk in ("value", "p50", "p99")}
                  for k, v in modes.items()},
        "end_to_end": e2e,
        "speedup_vs_plain_fp64": round(mixed_b["value"] / fp64["value"], 3),
        "accuracy_mixed_vs_fp64_over_diag": accuracy,
        "mixed_batch_inner_cycles_2": inner2,
        "mesh": {"vertices": modes["mixed"]["nv"], "fp64_vertices": fp64["nv"]},
        "roofline": {
            "kernel": f"kb_march<4, 33>: two red-black iterations (four colour sweeps) of the "
                      f"finest level of all {args.batch} instances of the batch in one "
                      "z-marching pass (one launch)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": round(alg_bytes),
            "algorithmic_bytes_rule": "8 n^3 words of 4 B per instance: two red-black "
                                      "iterations of 4 words per node (read x, rhs, screen, "
                                      "write x), fused or not",
            "us_per_launch": round(dom["ms"] * 1e3, 2),
            "loaded_bytes": round(dom["bytes"]),
            "loaded_frac": round(dom["gbs"] / HBM_PEAK_GBS, 4),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
            "traffic_GBps": (round(traffic / (dom["ms"] * 1e-3) / 1e9, 1)
                             if traffic is not None else None),
            "limiter": "VALU: the pass recomputes halo rows and moves about half of the "
                       "algorithmic bytes, so frac above 1 reads as the fusion's saving and "
                       "traffic_GBps as the HBM rate it draws",
            "single_sweep": {"kernel": "kb_relax_group<float>, one colour (the unfused sweep)",
                             "us_per_launch": round(bkernels["rb_sweep_f32"]["ms"] * 1e3, 2),
                             "algorithmic_bytes": round(float(args.batch) * 2.0 * n ** 3 * 4.0),
                             "achieved": round(float(args.batch) * 2.0 * n ** 3 * 4.0 /
                                               (bkernels["rb_sweep_f32"]["ms"] * 1e-3) / 1e9, 1)},
        },
        "kernels": {k: {"us": round(v["ms"] * 1e3, 2), "GBps": round(v["gbs"], 1)}
                    for k, v in kernels.items()},
        "kernels_batch": {k: {"us": round(v["ms"] * 1e3, 2), "GBps": round(v["gbs"], 1),
                              "bytes": round(v["bytes"])} for k, v in bkernels.items()},
    }
    line.update(extras)
    if not args.no_cpu_baseline:
        line["cpu_baseline"], line["cpu_baseline_legs"] = cpu_baseline(points, normals, args.depth)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
