"""The request loop (vba_server_*) on the GPU: a mix of pose-only
refinements, local windows and matrix-free submaps submitted together, with
every result compared against the same request solved alone. Windows share
device batches but no arithmetic, and submaps run the single-problem
driver, so results must be bitwise equal to the solo solves."""
import numpy as np
import pytest

import vba
from vba import abi

pytestmark = pytest.mark.gpu


def mix():
    reqs = []
    cauchy = dict(robust_kernel=abi.KERNEL_CAUCHY, robust_scale=1.0)
    for i in range(6):
        p = vba.synthesize(abi.SYNTH_POSE_ONLY, num_landmarks=200 + 150 * i, seed=40 + i).problem
        reqs.append((p, vba.default_options(max_iterations=10, **cauchy)))
    for i in range(4):
        p = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=5 + 3 * i, num_landmarks=300 + 200 * i,
                           seed=60 + i).problem
        reqs.append((p, vba.default_options(max_iterations=10, **cauchy)))
    for i in range(2):
        p = vba.synthesize(abi.SYNTH_LOOP, num_cameras=60 + 20 * i, num_landmarks=3000, seed=80 + i,
                           scene_scale=16.0).problem
        reqs.append((p, vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE, max_iterations=8,
                                            max_pcg_iterations=100, pcg_relative_tolerance=1e-4,
                                            robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                                            precision=abi.PRECISION_MIXED,
                                            pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE)))
    return reqs


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
