"""Device solves of the full-size workloads against the oracle's stored LM
traces (tests/golden/traces, written in this repository by
tests/golden/make_traces.py from the CPU FP64 restatement of the reference
solver, lm.rs:162-1017).

The bar is the one the device must meet on every workload: the same
accept/reject sequence as the FP64 oracle, final cost within 1e-6 relative,
camera centres within 1e-6 of the scene extent (diagonal of the true
landmarks' bounding box). Each test prints the levels it actually reached.

- circle: 50 x 5,000 SIMPLE_RADIAL, dense, squared loss, 20 iterations.
- loop: 1,200 cameras, 400k landmarks, ~3.5M observations, Huber 1 px,
  matrix-free PCG (rtol 1e-4, max 100), 30 iterations, under both at-cap
  rules (keep the iterate; reject the step), FP64 and mixed precision.
- windows: all 256 local stereo windows at their full K (5-40) and L
  (300-8,000), solved as one device batch.
- city500: 500 cameras of the city scene with per-camera OPENCV intrinsics,
  10 % outliers, Huber 2 px, matrix-free, 25 iterations, FP64 and mixed.
- dense1500: a 1,500-camera loop through the dense reduced solve (9,000
  unknowns), 8 iterations.

A problem is regenerated from its seed and must hash to the digest stored
with its trace, so generator drift fails here instead of comparing
different problems."""
import os

import numpy as np
import pytest

import vba
from vba import abi, workloads

pytestmark = pytest.mark.gpu

TRACES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "traces")
COST_RTOL = 1e-6    # relative final cost vs the FP64 oracle
CENTRE_TOL = 1e-6   # camera-centre distance / scene extent


def load(name):
    with np.load(os.path.join(TRACES, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def centre_error(poses, centres):
    return float(np.linalg.norm(vba.camera_centres(poses) - centres, axis=1).max())


def decision_report(accepted, cost_after, tr_acc, tr_cost_before, tr_cost_after):
    """The iterations where the decisions differ, with the oracle's margin
    |cost_after - cost_before| / cost_before there."""
    out = []
    for i, (a, b) in enumerate(zip(accepted, tr_acc)):
        if bool(a) != bool(b):
            m = abs(tr_cost_after[i] - tr_cost_before[i]) / tr_cost_before[i]
            out.append(f"iteration {i}: device {'accept' if a else 'reject'}, oracle {'accept' if b else 'reject'}, "
                       f"oracle margin {m:.2e}, device cost_after {cost_after[i]:.17g}")
    return out


def check_single(label, out, tr, extent):
    assert out.status == abi.OK, out.message
    acc = [bool(a) for a in out.accepted]
    ref = [bool(a) for a in tr["accepted"]]
    diffs = decision_report(acc, [it["cost_after"] for it in out.iterations], ref, tr["cost_before"],
                            tr["cost_after"])
    assert len(acc) == len(ref) and not diffs, (label, len(acc), len(ref), diffs)
    rel = abs(out.result.final_cost - float(tr["final_cost"])) / float(tr["final_cost"])
    cen = centre_error(out.poses, tr["centres"]) / extent
    print(f"{label}: {len(acc)} iterations ({sum(acc)} accepted) as the oracle; rel. final cost {rel:.2e}, "
          f"max centre error {cen:.2e} x extent")
    assert rel <= COST_RTOL, (label, rel)
    assert cen <= CENTRE_TOL, (label, cen)
    return rel, cen


def regenerate(synth, tr):
    assert workloads.problem_digest(synth.problem) == str(tr["digest"]), "seeded generator drifted"
    assert abs(synth.scene_extent - float(tr["extent"])) == 0.0


def test_circle_full_size(device):
    tr = load("circle")
    s = workloads.circle_problem()
    regenerate(s, tr)
    out = device.optimize(s.problem, workloads.circle_options())
    check_single("circle fp64", out, tr, s.scene_extent)


@pytest.fixture(scope="module")
def loop():
    return workloads.loop_problem()


@pytest.mark.parametrize("policy", ["accept", "reject"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_loop_full_size(device, loop, policy, precision):
    tr = load(f"loop_{policy}")
    regenerate(loop, tr)
    pol = abi.PCG_MAXIT_ACCEPT_ITERATE if policy == "accept" else abi.PCG_MAXIT_REJECT
    prec = abi.PRECISION_FP64 if precision == "fp64" else abi.PRECISION_MIXED
    out = device.optimize(loop.problem, workloads.loop_options(prec, pol))
    check_single(f"loop {policy}-at-cap {precision}", out, tr, loop.scene_extent)
    pcg = [it["pcg_iterations"] for it in out.iterations]
    print(f"loop {policy} {precision}: PCG iterations device {sum(pcg)}, oracle {int(tr['pcg_iterations'].sum())}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_city_subproblem(device, precision):
    tr = load("city500")
    s = workloads.city_problem(500)
    regenerate(s, tr)
    prec = abi.PRECISION_FP64 if precision == "fp64" else abi.PRECISION_MIXED
    out = device.optimize(s.problem, workloads.city_options(prec))
    check_single(f"city500 {precision}", out, tr, s.scene_extent)


def test_dense_1500_cameras(device):
    """The dense reduced solve at 9,000 unknowns (a 1,500-camera loop),
    factored across the device."""
    tr = load("dense1500")
    s = workloads.dense_loop_problem()
    regenerate(s, tr)
    out = device.optimize(s.problem, workloads.dense_loop_options())
    check_single("dense 1500 cameras", out, tr, s.scene_extent)


def test_all_windows_full_size_in_one_batch(device):
    tr = load("windows")
    specs = [tuple(int(v) for v in row) for row in tr["specs"]]
    assert specs == workloads.window_specs()
    synths = [workloads.window_problem(sp) for sp in specs]
    for i, s in enumerate(synths):
        assert workloads.problem_digest(s.problem) == str(tr["digest"][i]), ("seeded generator drifted", i)
    opts = workloads.window_options()
    pb = device.prepare_batch([s.problem for s in synths], opts)
    try:
        out = pb.solve(opts, cap=16)
        assert out.status == abi.OK, out.message
        worst_rel = worst_cen = 0.0
        ip, cp = tr["iteration_ptr"], tr["camera_ptr"]
        for i, (s, w) in enumerate(zip(synths, out.windows)):
            assert w.status == abi.OK
            acc = [bool(a) for a in w.accepted]
            ref = [bool(a) for a in tr["accepted"][ip[i]:ip[i + 1]]]
            diffs = decision_report(acc, [it["cost_after"] for it in w.iterations], ref,
                                    tr["cost_before"][ip[i]:ip[i + 1]], tr["cost_after"][ip[i]:ip[i + 1]])
            assert acc == ref and not diffs, (i, specs[i], diffs)
            rel = abs(w.result.final_cost - tr["final_cost"][i]) / tr["final_cost"][i]
            poses, _ = pb.state(i)
            cen = centre_error(poses, tr["centres"][cp[i]:cp[i + 1]]) / tr["extent"][i]
            assert rel <= COST_RTOL, (i, rel)
            assert cen <= CENTRE_TOL, (i, cen)
            worst_rel, worst_cen = max(worst_rel, rel), max(worst_cen, cen)
        kmax = max(sp[0] for sp in specs)
        lmax = max(sp[1] for sp in specs)
        print(f"windows: {len(specs)} windows (K up to {kmax}, L up to {lmax}) match the oracle's decisions; "
              f"worst rel. final cost {worst_rel:.2e}, worst centre error {worst_cen:.2e} x extent")
    finally:
        pb.free()
