#!/usr/bin/env python3
"""Reconstructions per second and p50/p99 latency of depth-8 screened Poisson
reconstruction on one MI355X, FP32-mixed against FP64.

Headline workload (one "step" = one full reconstruction): 200,000 oriented points on a
torus (R = 0.6, r = 0.25), Gaussian position noise 0.002 x bbox diagonal, 5 degree normal
noise, normal lengths U(0.5, 1.5) with confidence weighting; depth 8 (257^3 grid),
reference defaults (point weight 4, 4 V-cycles of 3/3 smoothing, 40 coarsest iterations,
1.5 samples per node, scale 1.1). Points are resident in HBM before the timed region;
each reconstruction normalises them, splats, assembles, solves, computes the iso value and
extracts the mesh on the device and returns its counts to the host (the mesh stays in HBM).

Modes, each timed over --steps iterations after --warmup:
  fp64        the plain FP64 solve, one request at a time (job run + wait)
  mixed       FP32 V-cycles inside FP64 defect correction, one request at a time
  mixed_graph the same recorded once as a HIP graph and replayed
  mixed_batch --batch depth-8 reconstructions (different seeds) as one batched graph launch
`value` is the mixed_batch recon/s. Also reported: the 64-instance batch of uneven depths
and sizes (depths U{5..8}, points log-U[2k, 200k], sphere/torus/superquadric) with its
per-instance error against FP64, the depth-9 2M-point mixed/FP64 pair, a served request
stream, the dominant kernel's roofline and the CPU oracle baseline. Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mojo-poissonrecon_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the roofline kernel (kb_relax_group<float>, one colour of the
# 257^3 finest level of all instances of the batch), keyed by (depth, batch):
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from rocprofv3 counter runs (the sweep timing's plain
# launches), the factors calibrated on a plain elementwise copy
# (profiles/kb_relax_group_f32_depth8_pmc.txt).
ROOFLINE_TRAFFIC_BYTES = {(8, 8): 894_785_904}


def log(*args):
    print("[bench]", *args, file=sys.stderr, flush=True)


def hausdorff_bound(va: np.ndarray, vb: np.ndarray) -> float:
    """Symmetric max over vertices of the distance to the nearest vertex of the other mesh
    (an upper bound of the vertex-to-surface Hausdorff distance)."""
    from scipy.spatial import cKDTree

    if not len(va) or not len(vb):
        return 0.0 if len(va) == len(vb) else float("inf")
    return float(max(cKDTree(vb).query(va)[0].max(), cKDTree(va).query(vb)[0].max()))


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
    return round(float(np.percentile(lat, q)) * 1e3, 4)


def cpu_baseline(points, normals, depth: int, budget_s: float = 20.0) -> dict:
    import oracle

    times = []
    start = time.perf_counter()
    while not times or (time.perf_counter() - start + times[-1] < budget_s and len(times) < 3):
        t = time.perf_counter()
        oracle.reconstruct(points, normals, depth=depth, confidence=True)
        times.append(time.perf_counter() - t)
    return {
        "value": round(1.0 / min(times), 4),
        "unit": "recon/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(times)} full depth-{depth} reconstructions of the same {len(points)}-point "
                  f"torus by the serial FP64 C oracle (best of {len(times)})",
    }


def run_batch_mode(clouds, depths, precision, reps, kernel_reps=0, kernels=None):
    """Records the batched run as one graph, replays it; returns (seconds per launch, latencies,
    meshes). With `kernels` [(name, id)] it also times those batched kernels afterwards on the
    batch's stream (include/mpr_hip.h mpr_batch_time_kernel) into the dict `kernels_out`."""
    from mojo_poissonrecon.batch import Batch
    from mojo_poissonrecon.core import Job, mesh_from_job, prepare_reconstruct

    reqs = [prepare_reconstruct(p, n, depth=d, precision=precision) for (p, n), d in
            zip(clouds, depths)]
    jobs = [Job(r.depth, len(r.points), r.precision) for r in reqs]
    for j, r in zip(jobs, reqs):
        j.set_points(r.points, r.normals)
    b = Batch(jobs)
    b.capture([r.params for r in reqs])
    b.launch()
    b.wait()
    t0 = time.perf_counter()
    lat = []
    for _ in range(reps):
        s = time.perf_counter()
        b.launch()
        results = b.wait()
        lat.append(time.perf_counter() - s)
    total = time.perf_counter() - t0
    meshes = [mesh_from_job(j, res, len(r.points)) for j, res, r in zip(jobs, results, reqs)]
    timed_kernels = {}
    for name, kid in kernels or ():
        from mojo_poissonrecon import _lib
        ms, nbytes = _lib.ctypes.c_double(), _lib.ctypes.c_double()
        _lib.check(_lib.lib().mpr_batch_time_kernel(b.handle, kid, kernel_reps,
                                                     _lib.ctypes.byref(ms),
                                                     _lib.ctypes.byref(nbytes)),
                   "mpr_batch_time_kernel")
        timed_kernels[name] = dict(ms=ms.value, bytes=nbytes.value,
                                   gbs=nbytes.value / (ms.value * 1e-3) / 1e9)
    b.close()
    for j in jobs:
        j.close()
    if kernels:
        return total / reps, np.array(lat), meshes, timed_kernels
    return total / reps, np.array(lat), meshes


def depth9_run(steps: int):
    """Depth-9 workload: 2,000,000 superquadric points (exponents 0.3/0.3, half-extents
    0.8/0.6/0.4, sigma 0.001 diag, 3 deg normal noise, unit normals), mixed precision and
    all-FP64 single requests at the same depth, with both final FP64 residual norms."""
    from mojo_poissonrecon import _lib, shapes
    from mojo_poissonrecon.core import Job

    p, n = shapes.superquadric_workload(2_000_000, seed=0)
    params = Job.params(_lib.STAGE_NORMALIZE | _lib.STAGE_SOLVE | _lib.STAGE_EXTRACT
                        | _lib.STAGE_WORLD)
    out, meshes = {}, {}
    for name, prec in (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED)):
        job = Job(9, len(p), prec)
        job.set_points(p, n)

        def step(job=job):
            job.run(params)
            job.wait()

        total, lat = timed(step, steps, 1)
        r = job.wait()
        meshes[name] = job.fetch_mesh(r)[0]
        out[name] = {"recon_per_s": round(steps / total, 3), "p50_ms": pct(lat, 50),
                     "residual_norm": r.residual_norm, "iso_value": r.iso_value,
                     "vertices": int(r.vertex_count),
                     "footprint_gb": round(Job.footprint(9, len(p), prec) / 1e9, 2)}
        job.close()
    diag = float(np.linalg.norm(p.max(axis=0) - p.min(axis=0)))
    return {"workload": "superquadric 2,000,000 points, depth 9 (513^3), sigma 0.001 diag, "
                        "3 deg normal noise", **out,
            "hausdorff_mixed_vs_fp64_over_diag": hausdorff_bound(meshes["mixed"], meshes["fp64"])
            / diag,
            "residual_ratio_mixed_over_fp64": out["mixed"]["residual_norm"]
            / out["fp64"]["residual_norm"]}


def serve_stream(requests: int, qps: float, seed: int, warmup: int = 0):
    """Poisson arrivals at `qps`; each request a cloud of 5k-500k points (log-uniform) on a
    random shape, randomly rotated and scaled into [-1,1]^3, depth from {6,7,8,9} weighted
    4:3:2:1. A pool of pre-generated clouds is rotated per request to bound host time. The
    first `warmup` requests of the same stream fill the server's job pool (device buffers
    are allocated on first use) and are excluded from the latency percentiles."""
    from mojo_poissonrecon import shapes
    from mojo_poissonrecon.serve import ReconstructionServer

    rng = np.random.default_rng(seed)
    makers = (shapes.sphere, shapes.torus, shapes.superquadric)
    pool = []
    for i in range(24):
        npts = int(round(math.exp(rng.uniform(math.log(5e3), math.log(5e5)))))
        p, n = makers[i % 3](npts, seed=7000 + i)
        pool.append(shapes.perturb(p, n, seed=7000 + i, sigma=0.001, degrees=2.0))
    total = requests + warmup
    gaps = rng.exponential(1.0 / qps, total)
    arrivals = np.cumsum(gaps)
    depth_choice = rng.choice([6, 7, 8, 9], size=total, p=[0.4, 0.3, 0.2, 0.1])
    picks = rng.integers(0, len(pool), total)
    server = ReconstructionServer(max_inflight=16, max_batch=8, precision="mixed")
    done = []
    start = time.perf_counter()
    nxt = 0
    measured_from = None
    while len(done) < total:
        now = time.perf_counter() - start
        if measured_from is None and nxt >= warmup:
            measured_from = now
        while nxt < total and arrivals[nxt] <= now:
            p, n = pool[picks[nxt]]
            rot = shapes.random_rotation(rng)
            s = float(rng.uniform(0.5, 1.0))
            pts = p @ rot.T
            pts = pts / np.abs(pts).max() * s
            server.submit(pts, n @ rot.T, depth=int(depth_choice[nxt]), arrival=start + arrivals[nxt])
            nxt += 1
        done += server.step()
    wall = time.perf_counter() - start
    server.close()
    measured = [r for r in done if r.id >= warmup]
    lat = np.array([r.latency for r in measured])
    cold = np.array([r.latency for r in done if r.id < warmup]) if warmup else np.array([0.0])
    last = max(r.completed for r in measured) - start
    first = min(r.arrival for r in measured) - start
    return {
        "requests": requests, "warmup_requests": warmup, "qps_offered": qps,
        "recon_per_s": round(len(measured) / (last - first), 2),
        "warmup_p99_ms": pct(cold, 99),
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "batches": server.stats.batches,
        "peak_inflight": server.stats.peak_inflight,
        "peak_hbm_gb": round(server.stats.peak_bytes / 1e9, 2),
        "jobs_created": server.stats.jobs_created,
        "host_s": {"prepare": round(server.stats.host_prepare, 3),
                   "launch": round(server.stats.host_launch, 3),
                   "collect": round(server.stats.host_collect, 3), "wall": round(wall, 3)},
        "depth_mix": "6:7:8:9 = 4:3:2:1", "points": "5k-500k log-uniform",
    }


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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--points", type=int, default=200_000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-qps", type=float, default=100.0)
    ap.add_argument("--serve-warmup", type=int, default=100,
                    help="requests that fill the job pool before the measured stream")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served stream and print its line")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the 64-batch and serving runs")
    ap.add_argument("--kernel-reps", type=int, default=40)
    ap.add_argument("--no-autotune", action="store_true",
                    help="launch the batched kernels with the default workgroup size")
    ap.add_argument("--only", choices=("fp64", "mixed", "mixed_batch"),
                    help="time one single-request mode and print its line (for per-mode "
                         "kernel traces)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="one reconstruction per precision, then only the kernel timings "
                         "(a short program for hardware-counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(independent reconstructions shard one per GPU with no exchange)", file=sys.stderr)
        return 2

    import mojo_poissonrecon as mpr
    from mojo_poissonrecon import _build, _lib, shapes
    from mojo_poissonrecon.core import Job

    if not os.path.exists(_build.LIB):
        mpr.build()
    if not args.no_autotune:
        from mojo_poissonrecon import tune
        t0 = time.perf_counter()
        chosen = tune.autotune(max_depth=max(5, min(args.depth, 10)), reps=20)
        log(f"autotune {time.perf_counter() - t0:.2f} s:",
            " ".join(f"{k}/{p}/{s}={b}x{g}x{z}" for (k, p, s), (b, g, z) in sorted(chosen.items())))

    points, normals = shapes.torus_workload(args.points, seed=0)
    stages = (_lib.STAGE_NORMALIZE | _lib.STAGE_SOLVE | _lib.STAGE_EXTRACT | _lib.STAGE_WORLD)
    params = Job.params(stages, confidence=True)
    modes = {}
    meshes = {}
    jobs = {}
    if args.serve_only:
        line = serve_stream(args.serve_requests, args.serve_qps, seed=2, warmup=args.serve_warmup)
        print(json.dumps({"mode": "serve", **line}), flush=True)
        return 0
    if args.only == "mixed_batch":
        bclouds = [shapes.torus_workload(args.points, seed=s) for s in range(args.batch)]
        per_launch, blat, _, bk = run_batch_mode(
            bclouds, [args.depth] * args.batch, "mixed", args.steps,
            kernel_reps=args.kernel_reps,
            kernels=[("rb_sweep_f32", 7), ("residual_f32", 9), ("defect_f64_f32", 13)])
        print(json.dumps({"mode": "mixed_batch", "batch": args.batch,
                          "recon_per_s": round(args.batch / per_launch, 3),
                          "p50_ms": pct(blat, 50), "steps": args.steps,
                          "kernels_batch": {k: {"us": round(v["ms"] * 1e3, 2),
                                                "GBps": round(v["gbs"], 1),
                                                "bytes": round(v["bytes"])}
                                            for k, v in bk.items()}}), flush=True)
        return 0
    runs = (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED))
    if args.only:
        runs = tuple(r for r in runs if r[0] == args.only)
    for name, prec in runs:
        job = Job(args.depth, len(points), prec)
        job.set_points(points, normals)

        def step(job=job):
            job.run(params)
            job.wait()

        total, lat = timed(step, args.steps, args.warmup)
        r = job.wait()
        meshes[name] = job.fetch_mesh(r)
        modes[name] = dict(value=args.steps / total, p50=pct(lat, 50), p99=pct(lat, 99),
                           nv=int(r.vertex_count), iso=r.iso_value, res=r.residual_norm)
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
    # graph replay of the mixed single request
    jm = jobs["mixed"]
    jm.capture(params)

    def gstep():
        jm.launch()
        jm.wait()

    total, lat = timed(gstep, args.steps, args.warmup)
    modes["mixed_graph"] = dict(value=args.steps / total, p50=pct(lat, 50), p99=pct(lat, 99))
    log(f"mixed_graph: {modes['mixed_graph']['value']:.2f} recon/s")
    # batched: --batch depth-8 reconstructions of the workload family in one graph launch
    bclouds = [shapes.torus_workload(args.points, seed=s) for s in range(args.batch)]
    per_launch, blat, bmeshes, bkernels = run_batch_mode(
        bclouds, [args.depth] * args.batch, "mixed", max(2, args.steps // 2),
        kernel_reps=args.kernel_reps,
        kernels=[("rb_sweep_f32", 7), ("residual_f32", 9), ("defect_f64_f32", 13)])
    modes["mixed_batch"] = dict(value=args.batch / per_launch, p50=pct(blat, 50),
                                p99=pct(blat, 99))
    log(f"mixed_batch x{args.batch}: {modes['mixed_batch']['value']:.2f} recon/s")

    diag = float(np.linalg.norm(points.max(axis=0) - points.min(axis=0)))
    hd = hausdorff_bound(meshes["mixed"][0], meshes["fp64"][0]) / diag

    kernels = time_kernels(jobs, args.kernel_reps)
    for job in jobs.values():
        job.close()
    # the dominant kernel of the headline workload: the batched FP32 colour sweep over the
    # finest level of all --batch instances, timed on the batch's stream
    dom = bkernels["rb_sweep_f32"]

    extras = {}
    if not args.no_extras:
        clouds, depths = shapes.uneven_batch_workload(args.instances, seed=1)
        t32, _, m32 = run_batch_mode(clouds, depths, "mixed", 3)
        t64, _, m64 = run_batch_mode(clouds, depths, "fp64", 3)
        errs = []
        for (p, _), a, b in zip(clouds, m32, m64):
            d = float(np.linalg.norm(p.max(axis=0) - p.min(axis=0)))
            errs.append(hausdorff_bound(a.vertices, b.vertices) / d)
        extras["batch64"] = {
            "instances": args.instances, "depths": "U{5,6,7,8}", "points": "2k-200k log-uniform",
            "mixed_recon_per_s": round(args.instances / t32, 2),
            "fp64_recon_per_s": round(args.instances / t64, 2),
            "max_hausdorff_mixed_vs_fp64_over_diag": max(errs),
            "launch_ms": round(t32 * 1e3, 3),
        }
        log("batch64:", extras["batch64"])
        extras["depth9"] = depth9_run(3)
        log("depth9:", extras["depth9"])
        extras["serve"] = serve_stream(args.serve_requests, args.serve_qps, seed=2,
                                       warmup=args.serve_warmup)
        log("serve:", extras["serve"])

    mixed_b, fp64 = modes["mixed_batch"], modes["fp64"]
    line = {
        "metric": "recon/s + p50/p99 latency, depth 8, FP32-mixed vs FP64 at fixed Hausdorff",
        "value": round(mixed_b["value"], 3),
        "unit": "recon/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / mixed_b["value"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycles in f64 defect correction (f64 splat, residual, iso)",
        "data": "synthetic",
        "config": {
            "workload": f"torus {len(points)} points, depth {args.depth} (257^3), sigma 0.002 diag, "
                        "5 deg normal noise, normal lengths U(0.5,1.5), confidence on",
            "mode": f"mixed precision, {args.batch} reconstructions per batched graph launch",
            "depth": args.depth,
            "points": len(points),
            "parallelism": "single GPU",
        },
        "p50_ms": mixed_b["p50"],
        "p99_ms": mixed_b["p99"],
        "modes": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                      for kk, vv in v.items() if kk in ("value", "p50", "p99")}
                  for k, v in modes.items()},
        "speedup_vs_plain_fp64": round(mixed_b["value"] / fp64["value"], 3),
        "hausdorff_mixed_vs_fp64_over_diag": hd,
        "mesh": {"vertices": modes["mixed"]["nv"], "fp64_vertices": fp64["nv"]},
        "roofline": {
            "kernel": f"kb_relax_group<float>: one colour of the finest level of all {args.batch} "
                      "instances of the batch (one launch)",
            "algorithmic_bytes": round(dom["bytes"]),
            "bound": "hbm",
            "achieved": round(dom["gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4),
            "traffic": ROOFLINE_TRAFFIC_BYTES.get((args.depth, args.batch)),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
        },
        "kernels": {k: {"us": round(v["ms"] * 1e3, 2), "GBps": round(v["gbs"], 1)}
                    for k, v in kernels.items()},
        "kernels_batch": {k: {"us": round(v["ms"] * 1e3, 2), "GBps": round(v["gbs"], 1),
                              "bytes": round(v["bytes"])} for k, v in bkernels.items()},
    }
    line.update(extras)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(points, normals, args.depth)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
