#!/usr/bin/env python3
"""Writes the oracle's LM traces of the full-size workloads as small fixtures
(tests/golden/traces/*.npz), for the device parity tests at those sizes.

Each trace holds the SHA-256 of the problem's bytes (so a change of the
seeded generator fails loudly instead of comparing different problems), the
per-iteration accept flags, cost before / after, post-update lambda and PCG
iterations, the initial and final cost, the final camera centres and the
scene extent. The oracle is the CPU FP64 restatement of the reference solver
(oracle/ba_oracle.cpp), run threaded (bitwise equal to serial).

Traces:
  circle        50 x 5,000 SIMPLE_RADIAL, dense, squared loss, 20 iterations
  loop_accept   1,200 x 400k loop, matrix-free FP64, 30 iterations, a PCG
                solve at the cap keeps its iterate (the GPU backend's rule)
  loop_reject   the same, a PCG solve at the cap fails the step (the CPU
                matrix-free entry's rule)
  windows       the 256 local stereo windows, dense, Cauchy 1 px, 10 iterations
  city500       a 500-camera sub-problem of the city scene (per-camera OPENCV,
                10 % outliers, Huber 2 px), matrix-free FP64, 25 iterations
  dense1500     a 1,500-camera loop (250 m, 60k landmarks), Huber 1 px, the
                dense reduced solve (9,000 unknowns), 8 iterations

Usage: python tests/golden/make_traces.py [--threads N] [name ...]
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "visloc-rs_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import vba  # noqa: E402
from vba import abi, workloads  # noqa: E402

OUT = os.path.join(HERE, "traces")


def trace_arrays(out, problem, extent):
    its = out.iterations
    return dict(
        status=np.int32(out.status),
        accepted=np.array([bool(i["step_accepted"]) for i in its], np.uint8),
        cost_before=np.array([i["cost_before"] for i in its], np.float64),
        cost_after=np.array([i["cost_after"] for i in its], np.float64),
        lam=np.array([i["lambda"] for i in its], np.float64),
        pcg_iterations=np.array([i["pcg_iterations"] for i in its], np.int32),
        initial_cost=np.float64(out.result.initial_cost),
        final_cost=np.float64(out.result.final_cost),
        centres=vba.camera_centres(out.poses),
        extent=np.float64(extent),
        num_observations=np.int64(problem.num_observations),
    )


def single(name, synth, options, threads):
    t0 = time.perf_counter()
    out = oracle.optimize(synth.problem, options, threads=threads)
    dt = time.perf_counter() - t0
    assert out.status == abi.OK, out.message
    arrs = trace_arrays(out, synth.problem, synth.scene_extent)
    arrs["digest"] = np.array(workloads.problem_digest(synth.problem))
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **arrs)
    print(f"{name}: N={synth.problem.num_observations} {len(out.iterations)} iterations "
          f"({int(arrs['accepted'].sum())} accepted), final cost {out.result.final_cost:.17g}, {dt:.1f} s",
          flush=True)


def windows(threads):
    specs = workloads.window_specs()
    opts = workloads.window_options()
    per = []
    t0 = time.perf_counter()
    for spec in specs:
        s = workloads.window_problem(spec)
        out = oracle.optimize(s.problem, opts, threads=threads)
        assert out.status == abi.OK, out.message
        a = trace_arrays(out, s.problem, s.scene_extent)
        a["digest"] = workloads.problem_digest(s.problem)
        per.append(a)
    dt = time.perf_counter() - t0

    def cat(key, dtype):
        return np.concatenate([np.atleast_1d(p[key]).astype(dtype) for p in per])

    it_ptr = np.concatenate([[0], np.cumsum([len(p["accepted"]) for p in per])]).astype(np.int64)
    cam_ptr = np.concatenate([[0], np.cumsum([len(p["centres"]) for p in per])]).astype(np.int64)
    np.savez_compressed(
        os.path.join(OUT, "windows.npz"),
        digest=np.array([p["digest"] for p in per]),
        specs=np.array(specs, np.int64),
        iteration_ptr=it_ptr, camera_ptr=cam_ptr,
        status=cat("status", np.int32), accepted=cat("accepted", np.uint8),
        cost_before=cat("cost_before", np.float64), cost_after=cat("cost_after", np.float64),
        lam=cat("lam", np.float64), pcg_iterations=cat("pcg_iterations", np.int32),
        initial_cost=cat("initial_cost", np.float64), final_cost=cat("final_cost", np.float64),
        centres=np.concatenate([p["centres"] for p in per]), extent=cat("extent", np.float64),
        num_observations=cat("num_observations", np.int64))
    print(f"windows: {len(specs)} windows, {int(sum(p['num_observations'] for p in per))} observations, "
          f"{dt:.1f} s", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("names", nargs="*",
                    default=["circle", "windows", "city500", "dense1500", "loop_accept", "loop_reject"])
    a = ap.parse_args()
    os.makedirs(OUT, exist_ok=True)
    for name in a.names:
        if name == "circle":
            single(name, workloads.circle_problem(), workloads.circle_options(), a.threads)
        elif name == "windows":
            windows(a.threads)
        elif name == "city500":
            single(name, workloads.city_problem(500), workloads.city_options(), a.threads)
        elif name == "dense1500":
            single(name, workloads.dense_loop_problem(), workloads.dense_loop_options(), a.threads)
        elif name in ("loop_accept", "loop_reject"):
            pol = abi.PCG_MAXIT_ACCEPT_ITERATE if name == "loop_accept" else abi.PCG_MAXIT_REJECT
            single(name, workloads.loop_problem(), workloads.loop_options(policy=pol), a.threads)
        else:
            raise SystemExit(f"unknown trace {name}")


if __name__ == "__main__":
    main()
