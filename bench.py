#!/usr/bin/env python3
"""Bundle-adjustment throughput on one MI355X.

Workload: the loop problem — 1,200 forward-facing pinhole cameras (f=800,
1024x768) on a closed 200 m loop, 400,000 landmarks in a 20 m corridor, each
seen by a run of consecutive cameras (~3.5M observations), 0.7 px noise,
Huber delta = 1 px, matrix-free block-Jacobi PCG (rtol 1e-4, max 100),
30 LM iterations. A step is one full solve (30 LM iterations) from the same
initial state with the problem resident in HBM (vba_prepare once, then
vba_solve_prepared per step).

value = observations x LM iterations executed (accepted + rejected), summed
over the timed steps, / wall seconds of the timed steps. The mixed-precision
solve is checked against the plain FP64 solve (final cost within 1e-6
relative) in the same run.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "visloc-rs_amd"))

import numpy as np  # noqa: E402

import vba  # noqa: E402
from vba import abi, workloads  # noqa: E402
from vba.workloads import loop_options, window_options, window_problems  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "BA obs/s + LM iters/s at ≤1e-6 rel cost vs FP64; served p50/p99 solve latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def cpu_threads():
    """Host threads for the threaded CPU baseline: the cores this process may
    use, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def as_fp64(options):
    o = vba.default_options(**{f: getattr(options, f) for f, _ in abi.Options._fields_})
    o.precision = abi.PRECISION_FP64
    return o


def cpu_sample(problems, options, label):
    """The CPU FP64 oracle (a restatement of the reference's Rust solver,
    bitwise equal serial and threaded) on a bounded sample: `problems` with
    `options`, timed serially (the reference default, parallel = false) and
    on cpu_threads() threads (parallel = true). Returns the timing record and
    the serial outputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    o = as_fp64(options)
    nthr = cpu_threads()
    t0 = time.perf_counter()
    outs = [oracle.optimize(p, o, threads=1) for p in problems]
    serial = time.perf_counter() - t0
    t0 = time.perf_counter()
    outs_t = [oracle.optimize(p, o, threads=nthr) for p in problems]
    threaded = time.perf_counter() - t0
    same = all(a.result.final_cost == b.result.final_cost and a.iterations == b.iterations
               for a, b in zip(outs, outs_t))
    obs = sum(p.num_observations * len(r.iterations) for p, r in zip(problems, outs))
    lm = sum(len(r.iterations) for r in outs)
    pcg = sum(max(0, it["pcg_iterations"]) for r in outs for it in r.iterations)
    return {
        "value": obs / serial,
        "unit": "obs/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{label}: {lm} LM iterations ({pcg} PCG iterations), FP64, "
                  f"{serial:.1f} s serial, {threaded:.1f} s on {nthr} threads (bitwise equal: {same}); "
                  f"{cpu_model()}, nproc {os.cpu_count()}",
        "seconds": serial,
        "threaded": {"value": obs / threaded, "cores": nthr, "seconds": threaded, "bitwise_equal_to_serial": same},
        "lm_iterations": lm,
        "pcg_iterations": pcg,
    }, outs


def device_sample(dev, problems, options, reps=3):
    """The device on the same sample as cpu_sample (same options, so the
    same LM iterations), for a like-for-like ratio: best of `reps` solves."""
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        for p in problems:
            dev.optimize(p, options)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best


def bench_windows(args):
    t0 = time.perf_counter()
    probs = window_problems(args.windows)
    n_obs = [p.num_observations for p in probs]
    log(f"bench: {len(probs)} windows, {sum(n_obs)} observations (generated in {time.perf_counter() - t0:.1f} s)")
    opts = window_options()
    dev = vba.Device(0)
    t0 = time.perf_counter()
    pb = dev.prepare_batch(probs, opts)
    log(f"bench: batch prepared in {time.perf_counter() - t0:.2f} s")
    for _ in range(args.warmup):
        out = pb.solve(opts, cap=16)
        if out.status != abi.OK:
            raise SystemExit(f"warmup batch failed: {out.status} {out.message}")
    obs_iters = 0
    lm_iters = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = pb.solve(opts, cap=16)
        if out.status != abi.OK:
            raise SystemExit(f"timed batch failed: {out.status} {out.message}")
        for n, w in zip(n_obs, out.windows):
            obs_iters += n * w.result.num_iterations
            lm_iters += w.result.num_iterations
    wall = time.perf_counter() - t0
    dev_s = out.windows[0].result.device_seconds
    log(f"bench: {args.steps} batch solves in {wall:.3f} s (device {dev_s:.3f} s per solve), {lm_iters} LM iterations")
    # parity sample: a few windows solved alone must match their batch results exactly
    for i in (0, len(probs) // 2, len(probs) - 1):
        alone = dev.optimize(probs[i], opts)
        assert alone.result.final_cost == out.windows[i].result.final_cost, i
    pb.free()
    dev.close()
    result = {
        "metric": METRIC,
        "value": obs_iters / wall,
        "unit": "obs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(probs)} local-BA stereo windows, K in [5,40], L in [300,8000], Cauchy 1 px, "
                               "dense reduced solve, 10 LM iterations, one batch", "observations": sum(n_obs)},
        "lm_iters_per_s": lm_iters / wall,
        "windows_per_s": len(probs) * args.steps / wall,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        sample = probs[:4]
        t1 = time.perf_counter()
        obs = 0
        for p in sample:
            r = oracle.optimize(p, opts, threads=1)
            obs += p.num_observations * len(r.iterations)
        dt = time.perf_counter() - t1
        result["cpu_baseline"] = {"value": obs / dt, "unit": "obs/s", "cores": 1, "kind": "port",
                                  "sample": f"the first {len(sample)} windows, 10 LM iterations each, FP64 dense, "
                                            f"serial; {dt:.1f} s on {cpu_model()}"}
    print(json.dumps(result), flush=True)


REDUCED_PUBLISHED_MS = 3.4  # GPU ms per PCG iteration at 3.4M camera-pair edges (docs/gpu_bundle_adjustment.md:97-107)


def reduced_system(n=5822, reach=620, seed=13):
    """A reduced camera system of SmallCity's size for the rig solver's
    iterative solve: n nodes in capture order, each coupled to its +-reach
    neighbours (3.42M edges at the defaults), FP32 (row b, column a) blocks
    whose size falls with capture distance, and diagonal blocks that make the
    matrix diagonally dominant, hence SPD."""
    rng = np.random.default_rng(seed)
    a = np.concatenate([np.arange(0, n - d, dtype=np.uint32) for d in range(1, reach + 1)])
    d = np.concatenate([np.full(n - d, d, dtype=np.uint32) for d in range(1, reach + 1)])
    order = np.lexsort((d, a))
    a, d = a[order], d[order]
    edges = np.stack([a, a + d], axis=1).astype(np.uint32)
    off = rng.standard_normal((len(edges), 6, 6), dtype=np.float32)
    off *= (0.05 / (1.0 + d / 32.0)).astype(np.float32)[:, None, None]
    row = np.abs(off).sum(axis=(1, 2)).astype(np.float64)
    weight = np.bincount(edges[:, 0], weights=row, minlength=n) + np.bincount(edges[:, 1], weights=row, minlength=n)
    g = rng.standard_normal((n, 6, 6)) * 0.1
    diag = np.einsum("nij,nkj->nik", g, g) + np.eye(6)[None] * (weight + 1.0)[:, None, None]
    rhs = rng.standard_normal(6 * n)
    return diag, edges, off, rhs


def cluster_runs(diag, edges, off, k):
    """Dense inverses of the matrix restricted to runs of k consecutive nodes
    (the rig solver's cluster_preconditioner), from the edges inside runs."""
    n = len(diag)
    inside = (edges[:, 0] // k) == (edges[:, 1] // k)
    mats = []
    for s in range(0, n, k):
        m = min(k, n - s)
        M = np.zeros((6 * m, 6 * m))
        for i in range(m):
            M[6 * i:6 * i + 6, 6 * i:6 * i + 6] = diag[s + i]
        mats.append((s, M))
    for (ea, eb), blk in zip(edges[inside], off[inside].astype(np.float64)):
        M = mats[ea // k][1]
        ia, ib = ea % k, eb % k
        M[6 * ib:6 * ib + 6, 6 * ia:6 * ia + 6] = blk
        M[6 * ia:6 * ia + 6, 6 * ib:6 * ib + 6] = blk.T
    return [(s, np.linalg.inv(M)) for s, M in mats]


def bench_reduced(args):
    """The ReducedSolveAccelerator path at the reference's published size:
    ms per PCG iteration, from two solves with different fixed iteration
    counts (tolerance 0), so the per-solve upload and setup cancel."""
    t0 = time.perf_counter()
    diag, edges, off, rhs = reduced_system()
    clusters = cluster_runs(diag, edges, off, 32)
    n, E = len(diag), len(edges)
    log(f"bench: reduced system {n} nodes, {E} edges, runs of 32 (built in {time.perf_counter() - t0:.1f} s)")
    dev = vba.Device(0)
    k1, k2 = 5, 5 + 20 * max(1, args.steps)

    def solve(k):
        t = time.perf_counter()
        out = dev.reduced_solve(diag, edges, off, rhs, 0.0, k, clusters=clusters, structure_id=1)
        if out is None or out[1] != k:
            raise SystemExit(f"reduced solve failed: {dev.last_error()}")
        return time.perf_counter() - t

    for _ in range(args.warmup):
        solve(k1)
    t1 = min(solve(k1) for _ in range(2))
    t2 = min(solve(k2) for _ in range(2))
    per_it = (t2 - t1) / (k2 - k1)
    dev.close()
    # algorithmic bytes per iteration: every FP32 block once, its edge ids, the
    # diagonal blocks, the cluster inverses and the PCG vectors
    cl_bytes = sum(inv.size for _, inv in clusters) * 8.0
    algo = E * (144.0 + 8.0) + n * 288.0 + cl_bytes + n * 6 * 8.0 * 10
    result = {
        "metric": "reduced-camera PCG ms per iteration at 3.4M camera-pair edges",
        "value": 1000.0 * per_it,
        "unit": "ms/iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t2,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": 1000.0 * per_it / REDUCED_PUBLISHED_MS,
        "dtype": "f32 blocks, f64 arithmetic",
        "data": "synthetic",
        "config": {"workload": f"reduced camera system of {n} nodes (SmallCity's frame count), {E} camera-pair "
                               "edges (+-620 capture-order neighbours), FP32 blocks, runs of 32 nodes, "
                               "fixed iteration counts", "nodes": n, "edges": E},
        "baseline_note": "vs_baseline divides by the reference's published 3.4 ms per PCG iteration at 3.4M edges "
                         "on an NVIDIA L4 (docs/gpu_bundle_adjustment.md:97-107)",
        "roofline": {"kernel": "whole PCG iteration", "bound": "hbm", "achieved": round(algo / per_it / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(algo / per_it / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "algorithmic_bytes_per_iteration": algo},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import reduced_pcg
        offd = off.astype(np.float64)
        ts = []
        for k in (1, 3):
            t = time.perf_counter()
            reduced_pcg.pcg(diag, edges, offd, rhs, 0.0, k, clusters=clusters)
            ts.append(time.perf_counter() - t)
        cpu_it = (ts[1] - ts[0]) / 2
        result["cpu_baseline"] = {"value": 1000.0 * cpu_it, "unit": "ms/iteration", "cores": 1, "kind": "port",
                                  "sample": "the oracle's numpy PCG (oracle/reduced_pcg.py) on the same system, "
                                            f"3 against 1 iterations; {cpu_model()}"}
    print(json.dumps(result), flush=True)


def serve_pools(seed=11, n_pose=48, n_win=32, n_sub=6):
    """Request pools of the served mixed stream: pose-only refinements (1
    camera, 200-2000 fixed points, 1 px noise, 5% outliers), local windows
    (as the windows workload) and submaps (100-300 cameras, 20k-120k
    landmarks on a loop trajectory)."""
    rng = np.random.default_rng(seed)
    dense = vba.default_options(linear_solver=abi.SOLVER_DENSE, max_iterations=10, robust_kernel=abi.KERNEL_CAUCHY,
                                robust_scale=1.0)
    sub_opts = loop_options(abi.PRECISION_MIXED)
    pose = [vba.synthesize(abi.SYNTH_POSE_ONLY, num_landmarks=int(rng.integers(200, 2001)), seed=5000 + i).problem
            for i in range(n_pose)]
    wins = window_problems(n_win, seed=seed)
    subs = []
    for i in range(n_sub):
        cams = int(rng.integers(100, 301))
        subs.append(vba.synthesize(abi.SYNTH_LOOP, num_cameras=cams, num_landmarks=int(rng.integers(20000, 120001)),
                                   seed=7000 + i, scene_scale=200.0 * cams / 1200.0).problem)
    return {"pose": (pose, dense), "window": (wins, dense), "submap": (subs, sub_opts)}


def serve_once(srv, pools, rate, seconds, seed=3):
    """One Poisson arrival window at `rate` req/s for `seconds` against a
    running server; returns per-request (kind, status, latency, obs*iters,
    arrival) and the wall time until the last result."""
    import threading
    rng = np.random.default_rng(seed)
    n_req = rng.poisson(rate * seconds)
    arrivals = np.sort(rng.uniform(0.0, seconds, n_req))
    kinds = rng.choice(["pose", "window", "submap"], size=n_req, p=[0.6, 0.3, 0.1])
    picks = [int(rng.integers(0, len(pools[k][0]))) for k in kinds]
    tickets = []
    lock = threading.Lock()
    done = []
    submitting = [True]

    def waiter():
        i = 0
        while True:
            with lock:
                if i >= len(tickets):
                    if not submitting[0]:
                        return
                    t = None
                else:
                    t, kind, p, at = tickets[i]
            if t is None:
                time.sleep(0.0005)
                continue
            r = srv.wait(t)
            done.append((kind, r.status, r.latency, p.num_observations * r.result.num_iterations, at))
            i += 1

    th = threading.Thread(target=waiter)
    th.start()
    start = time.perf_counter()
    for at, kind, pick in zip(arrivals, kinds, picks):
        now = time.perf_counter() - start
        if at > now:
            time.sleep(at - now)
        prob, opts = pools[kind][0][pick], pools[kind][1]
        t = srv.submit(prob, opts, borrow=True)
        with lock:
            tickets.append((t, kind, prob, time.perf_counter() - start))
    with lock:
        submitting[0] = False
    th.join()
    return done, time.perf_counter() - start


def serve_summary(done, wall, rate, seconds):
    lat = np.array([d[2] for d in done])
    finish = np.array([d[4] + d[2] for d in done])
    per_kind = {}
    for k in ("pose", "window", "submap"):
        lk = np.array([d[2] for d in done if d[0] == k])
        if len(lk):
            per_kind[k] = {"count": int(len(lk)), "p50_ms": round(1e3 * float(np.percentile(lk, 50)), 2),
                           "p99_ms": round(1e3 * float(np.percentile(lk, 99)), 2)}
    # requests still waiting when the arrival window closes, per second of window
    backlog = float(np.sum(finish > seconds)) / seconds
    drain = max(0.0, float(finish.max()) - seconds) if len(finish) else 0.0
    p99 = float(np.percentile(lat, 99)) if len(lat) else 0.0
    return {"rate": rate, "requests": len(done), "all_ok": all(d[1] == abi.OK for d in done),
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else 0.0, "p99_s": p99, "per_kind": per_kind,
            "solves_per_s": len(done) / wall, "served_obs_iters_per_s": sum(d[3] for d in done) / wall,
            "backlog_growth_req_per_s": backlog, "drain_s_after_window": drain,
            "bounded": bool(p99 <= 1.0 and drain <= 1.0)}


def bench_serve(args):
    """The served mixed stream: Poisson arrivals, 60/30/10 pose-only / local
    windows / submaps, against the request loop, at the stated rate and, with
    --serve-sweep, at lower rates to find the highest one whose p99 stays
    bounded (p99 <= 1 s and the queue drained within 1 s of the window)."""
    t0 = time.perf_counter()
    pools = serve_pools()
    log(f"bench: serve pools generated in {time.perf_counter() - t0:.1f} s")
    srv = vba.Server(window_streams=args.window_streams, submap_streams=args.submap_streams,
                     batch_wait_seconds=0.002)
    # warm the device paths once
    for k in ("pose", "window", "submap"):
        srv.wait(srv.submit(pools[k][0][0], pools[k][1], borrow=True))
    done, wall = serve_once(srv, pools, args.rate, args.serve_seconds)
    head = serve_summary(done, wall, args.rate, args.serve_seconds)
    log(f"bench: rate {args.rate:g}: {head}")
    sweep = []
    if args.serve_sweep:
        for r in [float(x) for x in args.serve_sweep.split(",")]:
            d, w = serve_once(srv, pools, r, args.serve_seconds, seed=int(r))
            sweep.append(serve_summary(d, w, r, args.serve_seconds))
            log(f"bench: rate {r:g}: p99 {sweep[-1]['p99_s']:.3f} s, bounded {sweep[-1]['bounded']}")
    counters = srv.counters()
    srv.close()
    # one-at-a-time plain FP64 solving of a sample, for comparison
    dev = vba.Device(0)
    solo = {}
    for k, cnt in (("pose", 12), ("window", 6), ("submap", 2)):
        probs, opts = pools[k]
        o = as_fp64(opts)
        ts = []
        for pr in probs[:cnt]:
            t1 = time.perf_counter()
            dev.optimize(pr, o)
            ts.append(time.perf_counter() - t1)
        solo[k] = float(np.mean(ts))
    dev.close()
    mix_mean = 0.6 * solo["pose"] + 0.3 * solo["window"] + 0.1 * solo["submap"]
    bounded = [x["rate"] for x in [head] + sweep if x["bounded"]]
    result = {
        "metric": METRIC,
        "value": head["p99_s"],
        "unit": "s (p99 latency)",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": 1000.0 * wall,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (windows, pose-only) / f32+f64 (submaps)",
        "data": "synthetic",
        "config": {"workload": f"served mixed stream: Poisson {args.rate:g} req/s for {args.serve_seconds:g} s, "
                               "60/30/10 pose-only / local windows / submaps, request pools of 48/32/6 problems, "
                               f"{args.window_streams} window and {args.submap_streams} submap lanes, "
                               "200 GB HBM budget", "requests": head["requests"]},
        **{k: head[k] for k in ("all_ok", "p50_s", "p99_s", "per_kind", "solves_per_s", "served_obs_iters_per_s",
                                "backlog_growth_req_per_s", "drain_s_after_window")},
        "offered_req_per_s": args.rate,
        "highest_bounded_rate": max(bounded) if bounded else None,
        "sweep": sweep,
        "counters": counters,
        "one_at_a_time_fp64": {"mean_s": solo, "solves_per_s": 1.0 / mix_mean},
    }
    print(json.dumps(result), flush=True)


def quat_rows(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def with_stereo_terms(p, rng, rect=0.3, gen=0.2, baseline=0.12):
    """The loop problem with 30 % of its observations turned rectified stereo
    (u_r from the current depth) and 20 % general stereo (a rotated,
    differently calibrated right camera), pixel noise 0.5 px."""
    import dataclasses
    fx, fy, cx, cy = (p.camera.params[i] for i in range(4))
    cam = vba.camera(abi.CAMERA_PINHOLE, [fx, fy, cx, cy], p.camera.width, p.camera.height)
    po, lo = np.argsort(p.pose_ids), np.argsort(p.landmark_ids)
    pk = po[np.searchsorted(p.pose_ids[po], p.obs_pose_ids)]
    lk = lo[np.searchsorted(p.landmark_ids[lo], p.obs_landmark_ids)]
    R = quat_rows(p.poses[:, :4])
    xc = np.einsum("nij,nj->ni", R[pk], p.landmarks[lk]) + p.poses[pk, 4:7]
    lr_q = np.array([0.99995, 0.005, -0.008, 0.003])
    lr_q /= np.linalg.norm(lr_q)
    lr_t = np.array([-baseline, 0.002, 0.001])
    xr = xc @ quat_rows(lr_q[None])[0].T + lr_t
    kind = rng.choice(3, size=len(pk), p=[1 - rect - gen, rect, gen])
    kind[(xc[:, 2] <= 0.1) | ((kind == 2) & (xr[:, 2] <= 0.1))] = 0
    noise = rng.normal(0.0, 0.5, (len(pk), 2))
    r, g = kind == 1, kind == 2
    st = np.column_stack([p.obs_xy[r], p.obs_xy[r, 0] - fx * baseline / xc[r, 2] + noise[r, 0]])
    gx = np.column_stack([p.obs_xy[g], 1.02 * fx * xr[g, 0] / xr[g, 2] + cx - 2 + noise[g, 0],
                          0.99 * fy * xr[g, 1] / xr[g, 2] + cy + 2 + noise[g, 1]])
    pair = vba.sensor(vba.camera(abi.CAMERA_PINHOLE, [1.02 * fx, 0.99 * fy, cx - 2, cy + 2]),
                      np.concatenate([lr_q, lr_t]))
    m = kind == 0
    return dataclasses.replace(
        p, camera=cam, obs_pose_ids=p.obs_pose_ids[m], obs_landmark_ids=p.obs_landmark_ids[m], obs_xy=p.obs_xy[m],
        stereo_pose_ids=p.obs_pose_ids[r], stereo_landmark_ids=p.obs_landmark_ids[r], stereo_obs=st,
        stereo_baseline=baseline, general_pose_ids=p.obs_pose_ids[g], general_landmark_ids=p.obs_landmark_ids[g],
        general_obs=gx, general_pair=np.zeros(int(g.sum()), np.uint32), pairs=[pair])


def bench_rows(args):
    """Throughput of the entry points beside the default path, each on a
    loop problem, one JSON line each: stereo observation kinds (mixed
    matrix-free), observation weights, GNC, the joint camera solve (dense)
    and the COLMAP binary round trip. Times are the caller's wall clock
    around one-shot entry points (layout, upload and solve) and the device
    time of the solve."""
    import tempfile
    rng = np.random.default_rng(0)
    dev = vba.Device(0)
    s = vba.synthesize(abi.SYNTH_LOOP, num_landmarks=args.landmarks // 4)
    p = s.problem
    mf = loop_options(abi.PRECISION_MIXED, abi.PCG_MAXIT_ACCEPT_ITERATE, abi.SOLVER_MATRIX_FREE)

    def line(name, problem, out, wall, extra=None):
        if out.status != abi.OK:
            raise SystemExit(f"{name}: status {out.status} {out.message}")
        it = out.result.num_iterations
        rec = {"metric": f"{name} observation-iterations/s", "value": problem.num_observations * it / wall,
               "unit": "obs*iter/s", "higher_is_better": True, "wall_s": wall,
               "device_s": out.result.device_seconds, "lm_iterations": it,
               "final_cost": out.result.final_cost, "observations": problem.num_observations,
               "config": {"workload": f"loop, {len(problem.pose_ids)} cameras, {len(problem.landmark_ids)} landmarks"}}
        rec.update(extra or {})
        print(json.dumps(rec), flush=True)

    st = with_stereo_terms(p, rng)
    dev.optimize(st, mf)  # warm
    t0 = time.perf_counter()
    out = dev.optimize(st, mf)
    line("stereo kinds (mixed, matrix-free)", st, out, time.perf_counter() - t0,
         {"rectified": len(st.stereo_pose_ids), "general": len(st.general_pose_ids)})

    w = rng.uniform(0.5, 1.0, p.num_observations)
    t0 = time.perf_counter()
    out = dev.optimize_weighted(p, mf, w)
    line("weighted (mixed, matrix-free)", p, out, time.perf_counter() - t0)

    q = p.copy()
    bad = rng.choice(len(q.obs_xy), size=len(q.obs_xy) // 50, replace=False)
    q.obs_xy[bad] += rng.uniform(-60.0, 60.0, (len(bad), 2))
    g = vba.default_gnc_options(kernel=abi.GNC_TRUNCATED_LEAST_SQUARES, c=3.0, max_outer=8, inner_iterations=3)
    t0 = time.perf_counter()
    st_, res, _, _, _ = dev.optimize_gnc(q, mf, g)
    wall = time.perf_counter() - t0
    if st_ != abi.OK:
        raise SystemExit(f"gnc: status {st_}")
    print(json.dumps({"metric": "GNC solve (TLS, 8 levels x 3 LM iterations, mixed matrix-free)", "value": wall,
                      "unit": "s", "higher_is_better": False, "outer_iterations": res.outer_iterations,
                      "observations": q.num_observations, "final_cost": res.final_cost,
                      "config": {"workload": f"loop with 2 % wrong matches, {len(q.pose_ids)} cameras"}}), flush=True)

    sj = vba.synthesize(abi.SYNTH_LOOP, num_cameras=400, num_landmarks=args.landmarks // 8)
    j = sj.problem.copy()
    fx, fy, cx, cy = (j.camera.params[i] for i in range(4))
    j.camera = vba.camera(abi.CAMERA_PINHOLE, [fx * 1.01, fy * 0.99, cx + 3, cy - 2], j.camera.width,
                          j.camera.height)
    oj = vba.default_options(refine_intrinsics=abi.REFINE_INTRINSICS | abi.REFINE_DISTORTION, max_iterations=10)
    t0 = time.perf_counter()
    out, cam = dev.optimize_intrinsics(j, oj)
    line("joint camera solve (dense, k = 6)", j, out, time.perf_counter() - t0,
         {"camera": list(cam.params[:cam.num_params])})

    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        vba.Model.from_problem(p).write_colmap(os.path.join(d, "m"), binary=True)
        t1 = time.perf_counter()
        model = vba.Model.read_colmap(os.path.join(d, "m"), binary=True)
        prob, _ = model.problem()
        t2 = time.perf_counter()
    print(json.dumps({"metric": "COLMAP binary write + read into a problem", "value": p.num_observations / (t2 - t0),
                      "unit": "obs/s", "higher_is_better": True, "write_s": t1 - t0, "read_s": t2 - t1,
                      "observations": int(prob.num_observations)}), flush=True)


TRACES = os.path.join(ROOT, "tests", "golden", "traces")
FP64_MFMA_PEAK_TFLOPS = 78.6


def load_trace(name):
    path = os.path.join(TRACES, f"{name}.npz")
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def parity_record(out, tr, extent, poses, idx=None):
    """Device solve against an oracle trace: same accept vector, relative
    final cost, largest camera-centre distance over the scene extent."""
    if idx is None:
        ref = [bool(a) for a in tr["accepted"]]
        ref_cost = float(tr["final_cost"])
        centres = tr["centres"]
    else:
        ip, cp = tr["iteration_ptr"], tr["camera_ptr"]
        ref = [bool(a) for a in tr["accepted"][ip[idx]:ip[idx + 1]]]
        ref_cost = float(tr["final_cost"][idx])
        centres = tr["centres"][cp[idx]:cp[idx + 1]]
    acc = [bool(it["step_accepted"]) for it in out.iterations]
    return {
        "same_accept_vector": acc == ref,
        "rel_final_cost": abs(out.result.final_cost - ref_cost) / ref_cost,
        "centre_err_over_extent": float(np.linalg.norm(vba.camera_centres(poses) - centres, axis=1).max()) / extent,
    }


def trace_parity(name, problem, out, extent, poses):
    tr = load_trace(name)
    if tr is None:
        return None
    if workloads.problem_digest(problem) != str(tr["digest"]):
        return {"error": "the problem's bytes differ from the trace's"}
    r = parity_record(out, tr, extent, poses)
    r["trace"] = name
    return r


def kernel_table(dev, names):
    out = {}
    for name in names:
        q = dev.profile_query(name)
        if q["launches"]:
            q["avg_us"] = 1e6 * q["seconds"] / q["launches"]
            out[name] = q
    return out


def dense_records(kern, assemble, factor):
    """Roofline of the dense assembly (HBM, algorithmic bytes) and FP64-MFMA
    utilisation of the dense factor (n^3 / 3 FLOPs per system)."""
    rec = {}
    if assemble in kern and kern[assemble]["seconds"] > 0:
        a = kern[assemble]
        gbs = a["bytes"] / a["seconds"] / 1e9
        rec["roofline"] = {"kernel": assemble, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                           "algorithmic_bytes_per_launch": a["bytes"] / a["launches"],
                           "avg_launch_us": round(a["avg_us"], 2)}
    if factor in kern and kern[factor]["seconds"] > 0:
        f = kern[factor]
        tf = f["bytes"] / f["seconds"] / 1e12
        rec["dense_factor_mfma"] = {"kernel": factor, "achieved_tflops": round(tf, 4), "peak": FP64_MFMA_PEAK_TFLOPS,
                                    "unit": "TFLOP/s (FP64)", "frac": round(tf / FP64_MFMA_PEAK_TFLOPS, 5),
                                    "flops_per_launch": f["bytes"] / f["launches"],
                                    "avg_launch_us": round(f["avg_us"], 2)}
    return rec


def extra_circle(dev, args, steps=5):
    """50 cameras on a circle, 5,000 landmarks, SIMPLE_RADIAL, every in-frame
    observation (250k), squared loss, dense Schur + Cholesky, 20 LM
    iterations (the solve converges earlier; iterations executed count)."""
    s = workloads.circle_problem()
    p = s.problem
    N = p.num_observations
    o = workloads.circle_options()
    prep = dev.prepare(p, o)
    prep.solve(o)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = prep.solve(o)
    wall = time.perf_counter() - t0
    poses = prep.state()[0]
    dev.profile(True, stride=1)
    prep.solve(o)
    dev.profile(False, reset=False)
    kern = kernel_table(dev, ["linearize", "cost", "dense_assemble", "dense_solve"])
    prep.free()
    lm = out.result.num_iterations * steps
    rec = {
        "workload": f"circle: 50 cameras, 5,000 landmarks, SIMPLE_RADIAL, {N} observations, squared loss, dense, "
                    "20 LM iterations max",
        "obs_per_s": N * lm / wall, "lm_iters_per_s": lm / wall, "ms_per_solve": 1e3 * wall / steps,
        "lm_iterations": out.result.num_iterations, "dtype": "f64",
        "vs_oracle_trace": trace_parity("circle", p, out, s.scene_extent, poses),
    }
    rec.update(dense_records(kern, "dense_assemble", "dense_solve"))
    if not args.no_cpu_baseline:
        cpu, _ = cpu_sample([p], o, "the full circle workload")
        cpu["device_seconds_same_sample"] = wall / steps
        rec["cpu_baseline"] = cpu
    return rec


def extra_windows(dev, args, steps=3):
    """The 256 local stereo windows (K in [5, 40], L in [300, 8000], Cauchy
    1 px, dense, 10 LM iterations) as one prepared batch replayed as a graph."""
    specs = workloads.window_specs()
    synths = [workloads.window_problem(sp) for sp in specs]
    probs = [x.problem for x in synths]
    n_obs = [x.num_observations for x in probs]
    o = workloads.window_options()
    pb = dev.prepare_batch(probs, o)
    pb.solve(o, cap=16)
    t0 = time.perf_counter()
    obs_iters = lm = 0
    for _ in range(steps):
        out = pb.solve(o, cap=16)
        for n, w in zip(n_obs, out.windows):
            obs_iters += n * w.result.num_iterations
            lm += w.result.num_iterations
    wall = time.perf_counter() - t0
    tr = load_trace("windows")
    parity = None
    if tr is not None:
        if [tuple(int(v) for v in r) for r in tr["specs"]] == specs and all(
                workloads.problem_digest(p) == str(d) for p, d in zip(probs, tr["digest"])):
            recs = [parity_record(w, tr, tr["extent"][i], pb.state(i)[0], idx=i) for i, w in enumerate(out.windows)]
            parity = {"windows": len(recs), "same_accept_vectors": sum(r["same_accept_vector"] for r in recs),
                      "max_rel_final_cost": max(r["rel_final_cost"] for r in recs),
                      "max_centre_err_over_extent": max(r["centre_err_over_extent"] for r in recs),
                      "trace": "windows"}
        else:
            parity = {"error": "the problems' bytes differ from the trace's"}
    # kernel times from one solve launched kernel by kernel (no graph)
    eager = vba.default_options(**{f: getattr(o, f) for f, _ in abi.Options._fields_})
    eager.flags = abi.FLAG_NO_GRAPH
    dev.profile(True, stride=1)
    pb.solve(eager, cap=16)
    dev.profile(False, reset=False)
    kern = kernel_table(dev, ["batch_dense_assemble", "batch_dense_solve"])
    pb.free()
    rec = {
        "workload": f"{len(probs)} local stereo windows, K in [5,40], L in [300,8000], {sum(n_obs)} observations, "
                    "Cauchy 1 px, dense, 10 LM iterations, one batch (HIP graph)",
        "obs_per_s": obs_iters / wall, "lm_iters_per_s": lm / wall, "ms_per_step": 1e3 * wall / steps,
        "dtype": "f64", "vs_oracle_trace": parity,
    }
    rec.update(dense_records(kern, "batch_dense_assemble", "batch_dense_solve"))
    if not args.no_cpu_baseline:
        sample = probs[:3]
        cpu, _ = cpu_sample(sample, o, f"the first {len(sample)} windows, 10 LM iterations each")
        rec["cpu_baseline"] = cpu
    return rec


def extra_city(dev, args, steps=2):
    """The city scene at full size: 5,500 cameras with per-camera OPENCV
    intrinsics, 2.5M landmarks (~20M observations), 10 % outliers, Huber
    2 px, matrix-free PCG (rtol 1e-4, max 100), 25 LM iterations, mixed
    precision, checked against the FP64 device solve."""
    t0 = time.perf_counter()
    s = workloads.city_problem()
    p = s.problem
    N = p.num_observations
    gen = time.perf_counter() - t0
    o = workloads.city_options(abi.PRECISION_MIXED)
    t0 = time.perf_counter()
    prep = dev.prepare(p, o)
    prep_s = time.perf_counter() - t0
    prep.solve(o)
    dev.profile(True, stride=args.profile_stride)
    t0 = time.perf_counter()
    lm = pcg = 0
    for _ in range(steps):
        out = prep.solve(o)
        lm += out.result.num_iterations
        pcg += out.result.total_pcg_iterations
    wall = time.perf_counter() - t0
    dev.profile(False, reset=False)
    kern = kernel_table(dev, ["linearize", "schur_apply", "pcg_pose", "pcg_step", "cost"])
    t1 = time.perf_counter()
    ref = prep.solve(workloads.city_options(abi.PRECISION_FP64))
    t_ref = time.perf_counter() - t1
    prep.free()
    roof = None
    if "schur_apply" in kern and kern["schur_apply"]["seconds"] > 0:
        k = kern["schur_apply"]
        gbs = k["bytes"] / k["seconds"] / 1e9
        roof = {"kernel": "schur_apply", "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                "algorithmic_bytes_per_launch": k["bytes"] / k["launches"], "avg_launch_us": round(k["avg_us"], 2)}
    rec = {
        "workload": f"city: 5,500 cameras, per-camera OPENCV, 2.5M landmarks, {N} observations, 10 % outliers, "
                    "Huber 2 px, matrix-free PCG rtol 1e-4 max 100, 25 LM iterations",
        "obs_per_s": N * lm / wall, "lm_iters_per_s": lm / wall, "ms_per_solve": 1e3 * wall / steps,
        "pcg_iters_per_solve": pcg / steps, "dtype": "f32+f64",
        "rel_cost_vs_fp64": abs(out.result.final_cost - ref.result.final_cost) / ref.result.final_cost,
        "same_accept_vector_as_fp64": out.accepted == ref.accepted,
        "fp64_obs_per_s": N * ref.result.num_iterations / t_ref,
        "generate_s": round(gen, 2), "prepare_s": round(prep_s, 2),
        "roofline": roof,
        "kernels": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kern.items()},
    }
    if not args.no_cpu_baseline:
        sub = workloads.city_problem(500).problem
        one = workloads.city_options(abi.PRECISION_FP64, max_iterations=1)
        cpu, _ = cpu_sample([sub], one, "the first LM iteration of the 500-camera city sub-problem "
                                        f"({sub.num_observations} observations)")
        rec["cpu_baseline"] = cpu
    return rec


def quat_matrix(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


def rms_reprojection(problem, poses, landmarks):
    """sqrt(robust_cost(None) / observations), the RMS reprojection error
    ba_gpu_replay.rs:19-29 reports: squared pixel residuals of the
    observations in front of their camera (problem.rs:497-526), for the
    PINHOLE camera a problem dump holds."""
    cam = problem.camera
    if cam.model != abi.CAMERA_PINHOLE:
        return None
    fx, fy, cx, cy = (cam.params[i] for i in range(4))
    pose_row = {int(i): k for k, i in enumerate(problem.pose_ids)}
    lm_row = {int(i): k for k, i in enumerate(problem.landmark_ids)}
    pi = np.fromiter((pose_row[int(i)] for i in problem.obs_pose_ids), np.int64, len(problem.obs_pose_ids))
    li = np.fromiter((lm_row[int(i)] for i in problem.obs_landmark_ids), np.int64, len(problem.obs_landmark_ids))
    R = quat_matrix(poses[:, :4])
    xc = np.einsum("nij,nj->ni", R[pi], landmarks[li]) + poses[pi, 4:7]
    front = xc[:, 2] > 0.0
    z = np.where(front, xc[:, 2], 1.0)
    r = np.stack([fx * xc[:, 0] / z + cx, fy * xc[:, 1] / z + cy], -1) - problem.obs_xy
    s = np.where(front, (r * r).sum(-1), 0.0)
    return float(np.sqrt(s.sum() / max(1, len(s))))


def replay_record(name, seconds, out, problem):
    acc = sum(1 for it in out.iterations if it["step_accepted"])
    rms = rms_reprojection(problem, out.poses, out.landmarks) if out.status == abi.OK else None
    log(f"{name}: {seconds:.2f}s, {len(out.iterations)} iterations ({acc} accepted), robust cost "
        f"{out.result.initial_cost:.6e} -> {out.result.final_cost:.6e}, rms reproj "
        f"{rms if rms is None else round(rms, 4)} px")
    return {"status": out.status, "seconds": seconds, "iterations": len(out.iterations), "accepted": acc,
            "initial_cost": out.result.initial_cost, "final_cost": out.result.final_cost, "rms_px": rms,
            "accept_vector": [bool(it["step_accepted"]) for it in out.iterations]}


def bench_replay(args):
    """crates/ba-gpu/examples/ba_gpu_replay.rs:31-141 on this library: read
    a VISLOC_SFM_BA_DUMP problem (vba_model_read_ba_problem), keep the local
    window of its last N poses with --local N, solve it on the device after a
    one-iteration warm-up, and, unless --no-cpu-baseline, with the CPU FP64
    oracle (the reference's `ba.optimize`) on the same options; print the
    reference's report lines and one JSON line. Options are the SfM's BA
    config (IncrementalSfmConfig::default().ba_config: BaConfig::default()
    with Huber delta = 3 px, incremental_sfm/mod.rs:648-651), with
    --iterations as the reference's flag; --replay-solver picks the reduced
    solve (dense is what BaConfig::default() asks for)."""
    model = vba.Model.read_ba_problem(args.replay)
    if args.local:
        model.keep_local_window(args.local)
    problem, weights = model.problem()
    model.close()
    fixed = int(np.count_nonzero(problem.pose_fixed))
    log(f"problem: {len(problem.pose_ids)} poses ({fixed} fixed), {len(problem.landmark_ids)} landmarks, "
        f"{problem.num_observations} observations (local window: {args.local or None})")
    kw = dict(robust_kernel=abi.KERNEL_HUBER, robust_scale=3.0)
    if args.iterations:
        kw["max_iterations"] = args.iterations
    if args.replay_solver == "matrix-free":
        # the GPU backend's PCG settings (crates/ba-gpu/src/solver.rs:36-48, 497-508)
        kw.update(linear_solver=abi.SOLVER_MATRIX_FREE, max_pcg_iterations=100, pcg_relative_tolerance=1e-4,
                  pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)
    elif args.replay_solver == "sparse":
        kw["linear_solver"] = abi.SOLVER_SPARSE
    options = vba.default_options(**kw)
    t0 = time.perf_counter()
    dev = vba.Device(0)
    log(f"device init {time.perf_counter() - t0:.2f}s")
    solve = (lambda o: dev.optimize_weighted(problem, o, weights)) if weights is not None else \
        (lambda o: dev.optimize(problem, o))
    warm = vba.default_options(**{f: getattr(options, f) for f, _ in abi.Options._fields_})
    warm.max_iterations = 1
    solve(warm)
    t0 = time.perf_counter()
    out = solve(options)
    device = replay_record("mi355x", time.perf_counter() - t0, out, problem)
    dev.close()
    rec = {"metric": "replay", "dump": os.path.basename(args.replay), "local": args.local or None,
           "poses": len(problem.pose_ids), "fixed_poses": fixed, "landmarks": len(problem.landmark_ids),
           "observations": problem.num_observations, "solver": args.replay_solver,
           "max_iterations": options.max_iterations, "device": device}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        nthr = cpu_threads()
        t0 = time.perf_counter()
        cpu = (oracle.optimize_weighted(problem, options, weights, threads=nthr) if weights is not None
               else oracle.optimize(problem, options, threads=nthr))
        rec["cpu"] = replay_record(f"cpu ({nthr} threads)", time.perf_counter() - t0, cpu, problem)
        rec["cpu"]["threads"] = nthr
        rec["same_accept_vector"] = rec["cpu"]["accept_vector"] == device["accept_vector"]
        if cpu.status == abi.OK and out.status == abi.OK:
            rec["rel_cost_vs_cpu"] = abs(out.result.final_cost - cpu.result.final_cost) / max(
                abs(cpu.result.final_cost), 1e-300)
            rec["speedup"] = rec["cpu"]["seconds"] / max(device["seconds"], 1e-12)
    print(json.dumps(rec), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="leave out the circle, windows and city results of the default line")
    ap.add_argument("--landmarks", type=int, default=400_000, help="landmark count (default: full size)")
    ap.add_argument("--pcg-policy", choices=["accept", "reject"], default="accept",
                    help="at the PCG iteration cap: use the iterate (GPU backend) or reject the step (CPU "
                         "matrix-free entry)")
    ap.add_argument("--workload", choices=["loop", "circle", "windows", "city", "serve", "reduced", "rows"],
                    default="loop",
                    help="loop: the 3.5M-observation single problem (default); windows: a batch of local windows; "
                         "serve: the request loop under a Poisson stream")
    ap.add_argument("--rate", type=float, default=2000.0, help="requests per second (--workload serve)")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window (--workload serve)")
    ap.add_argument("--serve-sweep", default="", help="comma-separated lower rates to try (--workload serve)")
    ap.add_argument("--window-streams", type=int, default=2, help="window lanes of the server (--workload serve)")
    ap.add_argument("--submap-streams", type=int, default=6, help="submap lanes of the server (--workload serve)")
    ap.add_argument("--windows", type=int, default=256, help="windows in the batch (--workload windows)")
    ap.add_argument("--profile-stride", type=int, default=32,
                    help="record kernel times around every n-th launch of each kernel family")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default PCG launch schedule")
    ap.add_argument("--replay", default="",
                    help="a problem dump (VISLOC_SFM_BA_DUMP) to solve on the device and the CPU, as ba_gpu_replay")
    ap.add_argument("--local", type=int, default=0, help="--replay: keep the last N poses variable")
    ap.add_argument("--iterations", type=int, default=0, help="--replay: LM iterations (default: BaConfig's)")
    ap.add_argument("--replay-solver", choices=["dense", "sparse", "matrix-free"], default="dense",
                    help="--replay: the reduced solve")
    ap.add_argument("--solver", choices=["matrix-free", "explicit"], default="matrix-free",
                    help="reduced camera PCG of the loop workload: implicit operator (the workload's own "
                         "setting) or the explicitly assembled block-sparse matrix")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)
    if args.replay:
        bench_replay(args)
        return
    if args.workload == "windows":
        bench_windows(args)
        return
    if args.workload in ("circle", "city"):
        dev = vba.Device(0)
        r = (extra_circle if args.workload == "circle" else extra_city)(dev, args)
        dev.close()
        print(json.dumps(r), flush=True)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "reduced":
        bench_reduced(args)
        return
    if args.workload == "rows":
        bench_rows(args)
        return

    t0 = time.perf_counter()
    s = vba.synthesize(abi.SYNTH_LOOP, num_landmarks=args.landmarks)
    problem = s.problem
    N = problem.num_observations
    log(f"bench: loop workload {len(problem.pose_ids)} cameras, {len(problem.landmark_ids)} landmarks, "
        f"{N} observations (generated in {time.perf_counter() - t0:.1f} s)")
    precision = abi.PRECISION_MIXED if args.precision == "mixed" else abi.PRECISION_FP64
    policy = abi.PCG_MAXIT_ACCEPT_ITERATE if args.pcg_policy == "accept" else abi.PCG_MAXIT_REJECT
    solver = abi.SOLVER_MATRIX_FREE if args.solver == "matrix-free" else abi.SOLVER_PCG_EXPLICIT
    opts = loop_options(precision, policy, solver)
    dev = vba.Device(0)
    t0 = time.perf_counter()
    prep = dev.prepare(problem, opts)
    log(f"bench: prepared (layout + upload) in {time.perf_counter() - t0:.2f} s")

    for w in range(args.warmup):
        out = prep.solve(opts)
        if out.status != abi.OK:
            raise SystemExit(f"warmup solve failed: {out.status} {out.message}")
        log(f"bench: warmup {w}: {out.result.num_iterations} LM iterations, {out.result.total_pcg_iterations} PCG, "
            f"cost {out.result.initial_cost:.6e} -> {out.result.final_cost:.6e}, "
            f"device {out.result.device_seconds:.3f} s")

    # The PCG launch schedule picked by timing the candidate schedules on this
    # problem (untimed; every candidate gives the same result bit for bit).
    tuning = None
    if not args.no_autotune:
        t1 = time.perf_counter()
        tuning = prep.autotune(opts, lm_iterations=6)
        log(f"bench: autotuned PCG schedule {tuning['tuning']} ({tuning['seconds'] * 1e3:.1f} ms vs "
            f"{tuning['default_seconds'] * 1e3:.1f} ms for the default, 6 LM iterations) in "
            f"{time.perf_counter() - t1:.2f} s")

    # HIP events on every 32nd launch of each kernel family: enough launches
    # for the averages, few enough that recording them does not slow the solve.
    # The PCG loop's kernels (product, pose, step) get the pair bound to their
    # dispatch, so it times the kernel alone; the others are bracketed on the
    # stream.
    dev.profile(True, stride=args.profile_stride)
    lm_iters = 0
    pcg_iters = 0
    final_cost = None
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = prep.solve(opts)
        if out.status != abi.OK:
            raise SystemExit(f"timed solve failed: {out.status} {out.message}")
        lm_iters += out.result.num_iterations
        pcg_iters += out.result.total_pcg_iterations
        final_cost = out.result.final_cost
    wall = time.perf_counter() - t0
    dev.profile(False, reset=False)
    timed_poses = prep.state()[0]
    accepted = sum(it["step_accepted"] for it in out.iterations)
    log(f"bench: {args.steps} steps in {wall:.3f} s; last solve {out.result.num_iterations} LM iterations "
        f"({accepted} accepted), {out.result.total_pcg_iterations} PCG iterations")

    kernels = {}
    for name in ("linearize", "schur_apply", "pcg_pose", "pcg_step", "cost", "explicit_assemble", "explicit_apply"):
        q = dev.profile_query(name)
        if q["launches"]:
            q["avg_us"] = 1e6 * q["seconds"] / q["launches"]
            q["achieved_gbs"] = q["bytes"] / q["seconds"] / 1e9 if q["seconds"] > 0 else 0.0
            q["bytes_per_launch"] = q["bytes"] / q["launches"]
            kernels[name] = q
            log(f"bench: kernel {name}: {q}")
    dominant = max((k for k in kernels if kernels[k]["bytes"] > 0), key=lambda k: kernels[k]["seconds"])
    dk = kernels[dominant]
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        entry = tr.get(f"{dominant}/{args.precision}")
        if entry:
            traffic = entry.get("hbm_bytes_per_launch")
    roofline = {
        "kernel": dominant,
        "bound": "hbm",
        "achieved": round(dk["achieved_gbs"], 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dk["achieved_gbs"] / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": dk["bytes_per_launch"],
        "avg_launch_us": round(dk["avg_us"], 2),
    }

    # Parity of the measured solve against the plain FP64 solve (untimed).
    ref_rel = None
    fp64_obs_per_s = None
    ref = None
    if precision == abi.PRECISION_MIXED:
        f64 = loop_options(abi.PRECISION_FP64, policy, solver)
        t1 = time.perf_counter()
        ref = prep.solve(f64)
        t_ref = time.perf_counter() - t1
        if ref.status == abi.OK:
            ref_rel = abs(final_cost - ref.result.final_cost) / abs(ref.result.final_cost)
            fp64_obs_per_s = N * ref.result.num_iterations / t_ref
        log(f"bench: fp64 solve final cost {ref.result.final_cost:.9e}, mixed {final_cost:.9e}, rel {ref_rel:.3e}, "
            f"fp64 {t_ref:.3f} s")

    value = N * lm_iters / wall
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "obs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if precision == abi.PRECISION_MIXED else "f64",
        "data": "synthetic",
        "config": {
            "workload": "loop: 1200 cameras on a 200 m closed loop, 400k landmarks, "
                        f"{N} observations, Huber 1 px, {args.solver} block-Jacobi PCG rtol 1e-4 max 100, "
                        "30 LM iterations",
            "observations": N,
            "precision": args.precision,
            "pcg_at_cap": args.pcg_policy,
        },
        "lm_iters_per_s": lm_iters / wall,
        "pcg_iters_per_s": pcg_iters / wall,
        "final_cost": final_cost,
        "rel_cost_vs_fp64": ref_rel,
        "fp64_obs_per_s": fp64_obs_per_s,
        "speedup_vs_fp64": (value / fp64_obs_per_s) if fp64_obs_per_s else None,
        "roofline": roofline,
        "kernels": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in kernels.items()},
        "pcg_schedule": tuning,
    }
    # Parity of the measured solve against the oracle's stored trace of the
    # same problem (tests/golden/traces, FP64, same at-cap rule), when present.
    if args.landmarks == 400_000:
        result["vs_oracle_trace"] = trace_parity(f"loop_{args.pcg_policy}", problem, out, s.scene_extent,
                                                 timed_poses)
        if ref is not None and ref.status == abi.OK:
            result["fp64_vs_oracle_trace"] = trace_parity(f"loop_{args.pcg_policy}", problem, ref, s.scene_extent,
                                                          prep.state()[0])

    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle, first LM iteration, serial and threaded) ...")
        one = loop_options(precision, policy, solver, max_iterations=1)
        cpu, _ = cpu_sample([problem], one, "the first LM iteration of the full loop workload")
        t1 = time.perf_counter()
        prep.solve(one)
        gpu_s = time.perf_counter() - t1
        cpu["device_seconds_same_sample"] = gpu_s
        cpu["device_speedup_same_sample"] = cpu["seconds"] / gpu_s
        result["cpu_baseline"] = cpu
    else:
        result["cpu_baseline"] = None
    prep.free()
    if not args.no_extra:
        result["extra"] = {}
        for name, fn in (("circle", extra_circle), ("windows", extra_windows), ("city", extra_city)):
            t1 = time.perf_counter()
            try:
                result["extra"][name] = fn(dev, args)
            except Exception as e:  # a failure is recorded, the headline line still prints
                result["extra"][name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"bench: extra {name} in {time.perf_counter() - t1:.1f} s: {result['extra'][name]}")
    dev.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
