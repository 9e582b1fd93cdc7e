"""The request loop (vba_server_*) on the GPU: a mix of pose-only
refinements, local windows and matrix-free submaps submitted together, with
every result compared against the same request solved alone. Windows share
device batches but no arithmetic, and submaps run the single-problem
driver, so results must be bitwise equal to the solo solves."""
import numpy as np
import pytest

import oracle
import vba
from vba import abi, workloads

pytestmark = pytest.mark.gpu


def mix(extents=None):
    """Pose-only refinements, local windows and mixed-precision matrix-free
    submaps; `extents` (a list) receives each request's scene extent."""
    reqs = []
    cauchy = dict(robust_kernel=abi.KERNEL_CAUCHY, robust_scale=1.0)
    synths = []
    for i in range(6):
        s = vba.synthesize(abi.SYNTH_POSE_ONLY, num_landmarks=200 + 150 * i, seed=40 + i)
        synths.append(s)
        reqs.append((s.problem, vba.default_options(max_iterations=10, **cauchy)))
    for i in range(4):
        s = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=5 + 3 * i, num_landmarks=300 + 200 * i,
                           seed=60 + i)
        synths.append(s)
        reqs.append((s.problem, vba.default_options(max_iterations=10, **cauchy)))
    for i in range(2):
        s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=60 + 20 * i, num_landmarks=3000, seed=80 + i,
                           scene_scale=16.0)
        synths.append(s)
        reqs.append((s.problem, vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=8,
                                                    max_pcg_iterations=100, pcg_relative_tolerance=1e-4,
                                                    robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                                                    precision=abi.PRECISION_MIXED,
                                                    pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)))
    if extents is not None:
        extents.extend(s.scene_extent for s in synths)
    return reqs


def as_fp64(o):
    f = vba.default_options(**{name: getattr(o, name) for name, _ in abi.Options._fields_})
    f.precision = abi.PRECISION_FP64
    return f


def test_served_results_match_the_oracle(device):
    """Every served result (pose-only, window and mixed-precision submap)
    against the CPU FP64 oracle: the same number of LM iterations, final cost
    within 1e-9 relative for the FP64 kinds and 1e-6 for the mixed submaps,
    camera centres within 1e-6 of the scene extent. The accept/reject
    sequence of each request's solo device solve equals the oracle's."""
    extents = []
    reqs = mix(extents)
    with vba.Server(window_streams=2, submap_streams=2, batch_wait_seconds=0.005) as srv:
        served = [srv.wait(t) for t in [srv.submit(p, o) for p, o in reqs]]
    for (p, o), got, ext in zip(reqs, served, extents):
        ref = oracle.optimize(p, as_fp64(o), threads=4)
        assert got.status == ref.status == abi.OK
        assert got.result.num_iterations == len(ref.iterations)
        mixed = o.precision == abi.PRECISION_MIXED
        rel = abs(got.result.final_cost - ref.result.final_cost) / ref.result.final_cost
        assert rel <= (1e-6 if mixed else 1e-9), rel
        d = np.linalg.norm(vba.camera_centres(got.poses) - vba.camera_centres(ref.poses), axis=1).max()
        assert d <= 1e-6 * ext, (d, ext)
        assert device.optimize(p, o).accepted == ref.accepted


def test_pose_only_requests_match_the_oracle(device):
    """The served stream's pose-only kind (one free camera over 200-2,000
    fixed points, 1 px noise, 5 % outliers, Cauchy 1 px, 10 iterations), solo
    and as one device batch, against the oracle."""
    probs = workloads.pose_only_problems(16, seed=29)
    opts = workloads.pose_only_options()
    batch = device.optimize_batch(probs, opts)
    assert batch.status == abi.OK
    for p, w in zip(probs, batch.windows):
        ref = oracle.optimize(p, opts)
        alone = device.optimize(p, opts)
        assert ref.status == alone.status == w.status == abi.OK
        assert alone.accepted == ref.accepted == w.accepted
        rel = abs(alone.result.final_cost - ref.result.final_cost) / ref.result.final_cost
        assert rel <= 1e-9, rel
        assert w.result.final_cost == alone.result.final_cost
        ext = float(np.linalg.norm(p.landmarks.max(0) - p.landmarks.min(0)))
        d = np.linalg.norm(vba.camera_centres(alone.poses) - vba.camera_centres(ref.poses), axis=1).max()
        assert d <= 1e-6 * ext, (d, ext)


def test_served_results_equal_solo(device):
    reqs = mix()
    with vba.Server(window_streams=2, submap_streams=2, batch_wait_seconds=0.005) as srv:
        tickets = [srv.submit(p, o, borrow=(i % 2 == 1)) for i, (p, o) in enumerate(reqs)]
        served = [srv.wait(t) for t in tickets]
        counters = srv.counters()
    assert counters["completed"] == len(reqs)
    assert counters["batched_windows"] == 10
    for (p, o), got in zip(reqs, served):
        alone = device.optimize(p, o)
        assert got.status == alone.status == abi.OK, (got.status, alone.message)
        assert got.result.final_cost == alone.result.final_cost
        assert got.result.num_iterations == alone.result.num_iterations
        assert np.array_equal(got.poses, alone.poses)
        assert np.array_equal(got.landmarks, alone.landmarks)
        assert got.latency > 0.0


def test_served_errors_and_pending(device):
    good = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=6, num_landmarks=300, seed=90).problem
    bad = good.copy()
    bad.rig_landmark_ids = bad.rig_landmark_ids.copy()
    bad.rig_landmark_ids[3] = 10 ** 9  # unknown landmark -> VBA_ERR_MISSING_LANDMARK
    o = vba.default_options(max_iterations=5)
    with vba.Server(window_streams=1, submap_streams=1) as srv:
        t_bad = srv.submit(bad, o)
        t_good = srv.submit(good, o)
        r_bad = srv.wait(t_bad)
        r_good = srv.wait(t_good)
    assert r_bad.status == abi.ERR_MISSING_LANDMARK
    assert r_bad.result.error_id == 10 ** 9
    assert r_good.status == abi.OK


def test_admission_keeps_batches_within_the_hbm_budget(device):
    """With a budget that holds two windows, no batch (nor two batches in
    flight together) holds more; a request larger than the budget still
    runs, by itself. Results stay those of the solo solves."""
    cauchy = dict(robust_kernel=abi.KERNEL_CAUCHY, robust_scale=1.0)
    p = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=8, num_landmarks=600, seed=70).problem
    o = vba.default_options(max_iterations=6, **cauchy)
    alone = device.optimize(p, o)
    with vba.Server(window_streams=2, submap_streams=1, batch_wait_seconds=0.01) as probe:
        est = probe.request_bytes(p, o)
    assert est > 0
    for budget, per_batch in ((int(2.5 * est), 2), (est // 2, 1)):
        with vba.Server(window_streams=2, submap_streams=1, batch_wait_seconds=0.01,
                        hbm_budget_bytes=budget) as srv:
            tickets = [srv.submit(p, o) for _ in range(8)]
            served = [srv.wait(t) for t in tickets]
            c = srv.counters()
        assert c["completed"] == 8 and c["batched_windows"] == 8
        assert c["batched_windows"] <= per_batch * c["batches"], (budget, c)
        for got in served:
            assert got.status == abi.OK
            assert got.result.final_cost == alone.result.final_cost
            assert np.array_equal(got.poses, alone.poses)
