"""Observation weights and graduated non-convexity in the oracle, pinned by
the reference's own tests: the GncState / estimate_scale_mad unit tests
(pipelines/slam/src/gnc.rs:288-513), the confidence-weight tests
(pipelines/slam/tests/bundle_adjustment.rs:1380-1453) and the GNC bundle
tests (pipelines/slam/tests/gnc_robust_ba.rs)."""
import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi

GM, TLS = abi.GNC_GEMAN_MCCLURE, abi.GNC_TRUNCATED_LEAST_SQUARES


def cfg(kernel, **kw):
    return oracle.default_gnc_options(kernel=kernel, **kw)


# ---- GncState (gnc.rs:288-410) --------------------------------------------------

def test_default_config_is_the_references():
    g = oracle.default_gnc_options()
    assert (g.kernel, g.c, g.anneal_factor, g.max_outer, g.inner_iterations, g.has_auto_scale,
            g.auto_scale_readapt) == (GM, 1.0, 1.4, 100, 5, 0, 0)


def test_gm_first_surrogate_is_nearly_least_squares():
    w, mu, _ = oracle.gnc_state_weights(cfg(GM), 100.0, 0, [100.0, 0.01])
    assert mu == 200.0
    assert abs(w[0] - (2.0 / 3.0) ** 2) < 1e-12 and w[1] > 0.999


def test_gm_terminal_recovers_geman_mcclure_weight():
    s = np.array([0.0, 0.5, 1.0, 4.0, 25.0])
    w, mu, terminal = oracle.gnc_state_weights(cfg(GM), 100.0, 200, s)
    assert terminal and abs(mu - 1.0) < 1e-12
    np.testing.assert_allclose(w, (1.0 / (s + 1.0)) ** 2, atol=1e-12)


def test_gm_weight_monotone_in_residual_and_in_annealing():
    s = np.arange(200) * 0.5
    w, _, _ = oracle.gnc_state_weights(cfg(GM), 50.0, 0, s)
    assert np.all(np.diff(w) <= 1e-15) and np.all((w >= 0) & (w <= 1))
    prev = np.inf
    for k in range(50):
        wk, _, terminal = oracle.gnc_state_weights(cfg(GM), 100.0, k, [25.0])
        assert wk[0] <= prev + 1e-15
        prev = wk[0]
        if terminal:
            break
    assert prev < 0.01


def test_tls_terminal_is_hard_threshold_and_band():
    w, _, terminal = oracle.gnc_state_weights(cfg(TLS), 100.0, 200, [0.9, 1.2])
    assert terminal and abs(w[0] - 1.0) < 1e-9 and w[1] < 1e-9
    s = np.arange(400) * 0.25
    w, mu, _ = oracle.gnc_state_weights(cfg(TLS), 100.0, 0, s)
    assert np.all((w >= 0) & (w <= 1))
    lo, hi = mu / (mu + 1.0), (mu + 1.0) / mu
    wb, _, _ = oracle.gnc_state_weights(cfg(TLS), 100.0, 0, [lo * 0.99, hi * 1.01])
    assert abs(wb[0] - 1.0) < 1e-12 and wb[1] < 1e-12
    # no residual above c^2: the problem starts terminal
    assert oracle.gnc_state_weights(cfg(TLS), 0.3, 0, [0.1])[2]


def test_anneal_is_deterministic():
    runs = [[oracle.gnc_state_weights(cfg(GM), 73.0, k, [1.0])[1] for k in range(40)] for _ in range(2)]
    assert runs[0] == runs[1]


# ---- estimate_scale_mad (gnc.rs:412-470) ------------------------------------------

def test_mad_scale_brackets_inliers_below_outliers():
    sq = [(0.9 + 0.2 * (i / 199.0)) ** 2 for i in range(200)] + [r * r for r in (20.0, 25.0, 30.0, 50.0)]
    c = oracle.estimate_scale_mad(sq, 3.5)
    assert 1.1 < c < 20.0


def test_mad_scale_is_robust_to_outlier_magnitude():
    inl = [(1.0 + 0.01 * i) ** 2 for i in range(100)]
    a = oracle.estimate_scale_mad(inl + [50.0 ** 2] * 20, 3.5)
    b = oracle.estimate_scale_mad(inl + [5000.0 ** 2] * 20, 3.5)
    assert abs(a - b) < 1e-9


def test_mad_scale_ignores_nan_and_needs_two_finite():
    assert oracle.estimate_scale_mad([], 3.5) is None
    assert oracle.estimate_scale_mad([np.nan, 4.0], 3.5) is None
    c = oracle.estimate_scale_mad([9.0, 16.0, np.nan], 2.0)
    assert abs(c - (3.5 + 2.0 * 1.4826 * 0.5)) < 1e-12


def test_mad_scale_grows_with_noise_level():
    tight = [(0.3 + 0.002 * i) ** 2 for i in range(100)]
    loose = [(2.4 + 0.016 * i) ** 2 for i in range(100)]
    assert oracle.estimate_scale_mad(loose, 3.5) > 5.0 * oracle.estimate_scale_mad(tight, 3.5)


# ---- observation weights (bundle_adjustment.rs:1380-1453) ----------------------------

def test_zero_weight_mutes_the_outlier():
    ba = scenes.over_determined_bundle(inject_outlier=True)
    w = np.ones(len(ba.obs_pose_ids))
    bad = [i for i in range(len(w)) if ba.obs_pose_ids[i] == 50 and ba.obs_landmark_ids[i] == 1]
    assert len(bad) == 1
    w[bad[0]] = 0.0
    out = oracle.optimize_weighted(ba, vba.default_options(), w)
    assert out.status == abi.OK
    assert out.result.final_cost < 1e-18
    centre = vba.camera_centres(out.poses)[list(ba.pose_ids).index(50)]
    assert np.linalg.norm(centre - [2.0, 0.0, 0.0]) < 1e-10
    # unit weights are the unweighted solve bit for bit
    ref = oracle.optimize(ba, vba.default_options())
    one = oracle.optimize_weighted(ba, vba.default_options(), np.ones(len(w)))
    assert one.result.final_cost == ref.result.final_cost and one.iterations == ref.iterations
    np.testing.assert_array_equal(one.poses, ref.poses)


def test_weights_scale_the_cost_linearly():
    ba = scenes.over_determined_bundle(inject_outlier=True)
    o = vba.default_options(max_iterations=1)
    one = oracle.optimize_weighted(ba, o, np.ones(len(ba.obs_pose_ids)))
    two = oracle.optimize_weighted(ba, o, 2.0 * np.ones(len(ba.obs_pose_ids)))
    assert two.result.initial_cost == 2.0 * one.result.initial_cost


def test_invalid_weights_are_rejected():
    ba = scenes.over_determined_bundle(inject_outlier=False)
    n = len(ba.obs_pose_ids)
    assert oracle.optimize_weighted(ba, vba.default_options(), [1.0]).status == abi.ERR_OBSERVATION_WEIGHT_COUNT
    w = np.ones(n)
    w[3] = np.nan
    out = oracle.optimize_weighted(ba, vba.default_options(), w)
    assert out.status == abi.ERR_INVALID_OBSERVATION_WEIGHT and out.result.error_id == 3
    w[3] = -1.0
    assert oracle.optimize_weighted(ba, vba.default_options(), w).status == abi.ERR_INVALID_OBSERVATION_WEIGHT
    o = vba.default_options(refine_intrinsics=1)
    assert oracle.optimize_weighted(ba, o, np.ones(n)).status == abi.ERR_WEIGHTS_WITH_INTRINSICS


# ---- GNC bundle adjustment (gnc_robust_ba.rs) --------------------------------------

def dense():
    return vba.default_options(linear_solver=abi.SOLVER_DENSE)


def centre_error(poses, truth):
    return np.linalg.norm(vba.camera_centres(poses) - vba.camera_centres(np.array([p for _, p in truth])),
                          axis=1).max()


def test_gm_identifies_every_outlier():
    outliers = [7, 18, 33, 41, 52]
    ba, _ = scenes.gnc_bundle(outliers, 70.0)
    st, res, _, _, w = oracle.optimize_gnc(ba, dense(), cfg(GM, c=5.0))
    assert st == abi.OK and res.observation_count == 60 and res.converged
    max_out = max(w[i] for i in outliers)
    min_in = min(w[i] for i in range(60) if i not in outliers)
    assert max_out < 1e-3 and min_in > 1000.0 * max_out
    assert res.inlier_cost < 0.05 * res.initial_cost


def test_tls_recovers_truth_where_least_squares_is_dragged_off():
    outliers = [7, 18, 33, 41, 52]
    ba, truth = scenes.gnc_bundle(outliers, 70.0)
    l2 = oracle.optimize(ba, dense())
    st, _, poses, _, _ = oracle.optimize_gnc(ba, dense(), cfg(TLS, c=5.0))
    assert st == abi.OK
    err = centre_error(poses, truth)
    assert err < centre_error(l2.poses, truth) and err < 1e-3


def test_tls_hard_rejects_outliers():
    outliers = [25, 38, 50]
    ba, truth = scenes.gnc_bundle(outliers, 90.0)
    st, _, poses, _, w = oracle.optimize_gnc(ba, dense(), cfg(TLS, c=5.0))
    assert st == abi.OK
    assert all(w[i] < 1e-6 for i in outliers)
    assert all(abs(w[i] - 1.0) < 1e-9 for i in range(60) if i not in outliers)
    assert centre_error(poses, truth) < 1e-4


def test_clean_scene_keeps_every_observation():
    ba, truth = scenes.gnc_bundle()
    st, _, poses, _, w = oracle.optimize_gnc(ba, dense(), cfg(TLS, c=5.0))
    assert st == abi.OK
    assert np.sum(w >= 0.5) == 60
    assert centre_error(poses, truth) < 1e-6


def test_gnc_propagates_validation_errors():
    empty = vba.BundleProblem(camera=scenes.pinhole_camera(), pose_ids=[], poses=np.zeros((0, 7)), pose_fixed=[],
                              landmark_ids=[], landmarks=np.zeros((0, 3)), landmark_fixed=[], obs_pose_ids=[],
                              obs_landmark_ids=[], obs_xy=np.zeros((0, 2)))
    assert oracle.optimize_gnc(empty, dense(), cfg(GM, c=5.0))[0] == abi.ERR_NO_POSES


def test_readapt_tightens_inlier_scale():
    outliers = [7, 18, 33, 41, 52]

    def run(readapt):
        ba, truth = scenes.gnc_bundle(outliers, 80.0)
        for k, pid in enumerate(ba.pose_ids):
            if int(pid) in (11, 12, 13):
                c = vba.camera_centres(ba.poses[k:k + 1])[0]
                ba.poses[k] = scenes.pose_at([c[0] + 0.12, c[1], c[2] + 0.05], 0.02)
        g = cfg(TLS, c=2.0, has_auto_scale=1, auto_scale=3.5, auto_scale_readapt=int(readapt))
        st, res, _, _, w = oracle.optimize_gnc(ba, dense(), g)
        assert st == abi.OK
        return res, w

    one, w_one = run(False)
    re, w_re = run(True)
    assert re.inlier_scale < one.inlier_scale
    assert sum(w_re[i] < 0.5 for i in outliers) >= sum(w_one[i] < 0.5 for i in outliers)
