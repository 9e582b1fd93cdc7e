"""Pins the CPU FP64 oracle (oracle/ba_oracle.cpp) before anything is checked
against it: the reference's known-answer tests, closed forms, finite
differences and a brute-force dense solve of the full normal system."""
import math
import os

import numpy as np
import pytest

import oracle
import scenes
import vba
from vba import abi, workloads


# ---------------------------------------------------------------------------
# implicit_schur.rs:893-912, 1117-1188 — hand_check_system fixture
# ---------------------------------------------------------------------------
def hand_check_system():
    A = np.zeros((6, 3))
    A[0, 0] = A[1, 1] = A[2, 2] = 1.0
    return oracle.System(
        diag=[np.eye(6) * 10.0, np.eye(6) * 10.0],
        b_p=np.arange(1, 13, dtype=float),
        h_ll=[np.eye(3) * 4.0],
        b_l=[[1.0, 2.0, 3.0]],
        cross=[[(0, A), (0, 2.0 * A), (1, -A)]],
    )


def synthetic_system():
    """implicit_schur.rs:866-891."""
    c0 = np.fromfunction(lambda r, c: 0.02 * (r + 1) * (c + 1), (6, 3))
    c1 = np.fromfunction(lambda r, c: 0.015 * (r + 2) * (c + 1), (6, 3))
    c2 = np.fromfunction(lambda r, c: 0.01 * (r + 3) * (c + 2), (6, 3))
    return oracle.System(
        diag=[np.eye(6) * 20.0, np.eye(6) * 22.0],
        b_p=0.03 * np.arange(1, 13),
        h_ll=[np.diag([7.0, 8.0, 9.0])],
        b_l=[[0.2, -0.1, 0.3]],
        cross=[[(0, c0), (0, c1), (1, c2)]],
    )


def full_normal(ns: oracle.System, lam: float):
    """The undamped-then-damped full (6P + 3L) normal matrix and -gradient,
    assembled directly (implicit_schur.rs:914-948)."""
    P, L = ns.P, ns.L
    n = 6 * P + 3 * L
    H = np.zeros((n, n))
    for p in range(P):
        H[6 * p:6 * p + 6, 6 * p:6 * p + 6] = ns.diag[p].reshape(6, 6) + lam * np.eye(6)
    for l in range(L):
        o = 6 * P + 3 * l
        H[o:o + 3, o:o + 3] = ns.h_ll[l].reshape(3, 3) + lam * np.eye(3)
        for e in range(ns.cross_ptr[l], ns.cross_ptr[l + 1]):
            p = ns.cross_pose[e]
            W = ns.cross[e].reshape(6, 3)
            H[6 * p:6 * p + 6, o:o + 3] += W
            H[o:o + 3, 6 * p:6 * p + 6] += W.T
    g = np.concatenate([ns.b_p, ns.b_l.reshape(-1)])
    return H, -g


@pytest.mark.parametrize("lam", [0.0, 0.5])
def test_hand_check_fixture_matches_closed_form(lam):
    ns = hand_check_system()
    S, rhs = ns.reduce_dense(lam)
    e_pose0 = 7.75 if lam == 0.0 else 8.5
    e_pose1 = 9.75 if lam == 0.0 else 10.277777777777779
    e_cross = 0.75 if lam == 0.0 else 0.6666666666666666
    for c in range(3):
        assert abs(S[c, c] - e_pose0) < 1e-12
        assert abs(S[6 + c, 6 + c] - e_pose1) < 1e-12
        assert abs(S[c, 6 + c] - e_cross) < 1e-12
        assert abs(S[6 + c, c] - e_cross) < 1e-12
    for c in range(3, 6):
        assert abs(S[c, c] - (10.0 + lam)) < 1e-12
        assert abs(S[6 + c, 6 + c] - (10.0 + lam)) < 1e-12
        assert S[c, 6 + c] == 0.0
    e_rhs0 = [-0.25, -0.5, -0.75] if lam == 0.0 else [-1 / 3, -2 / 3, -1.0]
    e_rhs1 = [-7.25, -8.5, -9.75] if lam == 0.0 else [-65 / 9, -76 / 9, -29 / 3]
    st, _, _, op_rhs = ns.operator(lam)
    assert st == 0
    for c in range(3):
        assert abs(rhs[c] - e_rhs0[c]) < 1e-12 and abs(op_rhs[c] - e_rhs0[c]) < 1e-12
        assert abs(rhs[6 + c] - e_rhs1[c]) < 1e-12 and abs(op_rhs[6 + c] - e_rhs1[c]) < 1e-12
    # brute force: the full normal system solved directly
    H, b = full_normal(ns, lam)
    full = np.linalg.solve(H, b)
    st, dp, dl = ns.solve_dense(lam)
    assert st == 0
    assert np.linalg.norm(dp - full[:12]) < 1e-10
    assert np.linalg.norm(dl - full[12:]) < 1e-10
    st, dp, dl, it, rn, tg = ns.solve_pcg(lam)
    assert st == 0 and rn <= tg
    assert np.linalg.norm(dp - full[:12]) < 1e-9
    assert np.linalg.norm(dl - full[12:]) < 1e-9


def test_implicit_operator_matches_explicit_schur_and_preconditioner():
    """implicit_schur.rs:952-1015, with the explicit Schur complement formed
    by numpy from the full normal matrix."""
    lam = 0.25
    ns = synthetic_system()
    H, b = full_normal(ns, lam)
    A, B, D = H[:12, :12], H[:12, 12:], H[12:, 12:]
    S = A - B @ np.linalg.inv(D) @ B.T
    x = 0.04 * np.arange(1, 13)
    st, y, _, rhs = ns.operator(lam, x=x)
    assert st == 0
    assert np.linalg.norm(S @ x - y) < 1e-12
    assert np.allclose(rhs, b[:12] - B @ np.linalg.solve(D, b[12:]), atol=1e-14)
    # pose-0 preconditioner block groups the two pose-0 cross blocks
    c01 = ns.cross[0].reshape(6, 3) + ns.cross[1].reshape(6, 3)
    M0 = np.eye(6) * (20.0 + lam) - c01 @ np.linalg.inv(np.diag([7.0, 8.0, 9.0]) + lam * np.eye(3)) @ c01.T
    probe = np.arange(1, 13, dtype=float)
    st, _, z, _ = ns.operator(lam, r=probe)
    assert np.linalg.norm(z[:6] - np.linalg.solve(M0, probe[:6])) < 1e-12
    # PCG step equals the direct step
    st, dp, dl, it, rn, tg = ns.solve_pcg(lam)
    full = np.linalg.solve(H, b)
    assert st == 0 and np.linalg.norm(dp - full[:12]) < 1e-9 and np.linalg.norm(dl - full[12:]) < 1e-9


def test_pcg_limits():
    ns = synthetic_system()
    st, *_ = ns.solve_pcg(0.25, max_iterations=0)
    assert st == -7  # MaxIterations: a zero budget cannot solve a nonzero rhs
    st, dp, dl, it, rn, tg = ns.solve_pcg(0.25, max_iterations=1, rtol=1e-12, atol=1e-12)
    assert st == -7 and it == 1


def random_system(seed, P=6, L=40):
    """A BA-shaped normal system J^T J (plus a small pose prior) from random
    2x6 / 2x3 Jacobian blocks, each landmark seen from 2-3 poses, so that PCG
    needs many iterations."""
    rng = np.random.default_rng(seed)
    diag = [0.5 * np.eye(6) for _ in range(P)]
    h_ll, b_l, cross = [], [], []
    for _ in range(L):
        hl = np.zeros((3, 3))
        entries = []
        for p in rng.choice(P, size=int(rng.integers(2, 4)), replace=False):
            jp, jl = rng.normal(size=(2, 6)), rng.normal(size=(2, 3))
            diag[p] = diag[p] + jp.T @ jp
            hl += jl.T @ jl
            entries.append((int(p), jp.T @ jl))
        h_ll.append(hl)
        b_l.append(rng.normal(size=3))
        cross.append(entries)
    return oracle.System(diag, rng.normal(size=6 * P), h_ll, b_l, cross)


def textbook_pcg(S, rhs, k):
    """k steps of preconditioned conjugate gradients from x = 0 with the
    6x6 block-diagonal of S as the preconditioner (Saad, Iterative Methods
    for Sparse Linear Systems, Algorithm 9.1)."""
    n = len(rhs)
    Minv = np.zeros_like(S)
    for p in range(0, n, 6):
        Minv[p:p + 6, p:p + 6] = np.linalg.inv(S[p:p + 6, p:p + 6])
    x = np.zeros(n)
    r = rhs.copy()
    z = Minv @ r
    d = z.copy()
    rho = r @ z
    for _ in range(k):
        q = S @ d
        alpha = rho / (d @ q)
        x = x + alpha * d
        r = r - alpha * q
        z = Minv @ r
        rho_next = r @ z
        d = z + (rho_next / rho) * d
        rho = rho_next
    return x


@pytest.mark.parametrize("k", [1, 3, 8])
def test_pcg_iterate_at_the_cap(k):
    """The at-cap rules of the PCG step: keeping the iterate (the GPU
    backend, ba-gpu/src/solver.rs:497-508) returns exactly the k-th textbook
    PCG iterate and its landmark completion; the CPU matrix-free rule
    (implicit_schur.rs:715-736) fails the step with MaxIterations."""
    lam = 0.1
    ns = random_system(11)
    H, b = full_normal(ns, lam)
    assert np.linalg.eigvalsh(H).min() > 0.0
    n = 6 * ns.P
    A, B, D = H[:n, :n], H[:n, n:], H[n:, n:]
    S = A - B @ np.linalg.solve(D, B.T)
    rhs = b[:n] - B @ np.linalg.solve(D, b[n:])
    xk = textbook_pcg(S, rhs, k)
    exact = np.linalg.solve(S, rhs)
    assert np.linalg.norm(xk - exact) > 1e-6 * np.linalg.norm(exact)  # the cap really cuts the solve
    st, dp, dl, it, rn, tg = ns.solve_pcg(lam, max_iterations=k, rtol=0.0, atol=0.0, accept_at_cap=True)
    assert st == 0 and it == k
    assert np.linalg.norm(dp - xk) <= 1e-10 * np.linalg.norm(xk)
    assert np.linalg.norm(dl - np.linalg.solve(D, b[n:] - B.T @ dp)) <= 1e-10 * np.linalg.norm(dl)
    assert abs(rn - np.linalg.norm(rhs - S @ dp)) <= 1e-9 * np.linalg.norm(rhs)
    if k > 1:  # not the previous iterate
        assert np.linalg.norm(dp - textbook_pcg(S, rhs, k - 1)) > 1e-8 * np.linalg.norm(xk)
    st, *_ , it, _, _ = ns.solve_pcg(lam, max_iterations=k, rtol=0.0, atol=0.0, accept_at_cap=False)
    assert st == -7 and it == k
    # converging before the cap is the same under both rules
    st_a, dp_a, *_ = ns.solve_pcg(lam, max_iterations=500, rtol=1e-13, atol=0.0, accept_at_cap=True)
    st_r, dp_r, *_ = ns.solve_pcg(lam, max_iterations=500, rtol=1e-13, atol=0.0, accept_at_cap=False)
    assert st_a == st_r == 0 and np.array_equal(dp_a, dp_r)
    assert np.linalg.norm(dp_a - exact) <= 1e-10 * np.linalg.norm(exact)


def lambda_rule_violations(iterations, o):
    """The damping bookkeeping of the LM loop, lm.rs:610-662 and 890-991:
    a rejected or singular step raises lambda to min(lambda * increase,
    max_lambda) and reports the raised value, and the loop stops once it
    reaches max_lambda (after a cost/feasibility rejection also when there is
    no initial lambda); an accepted step reports the lambda it was solved
    with and then lowers it to max(lambda * decrease, min_lambda) (held at 0
    without an initial lambda)."""
    bad = []
    lam = o.initial_lambda if o.has_initial_lambda else 0.0
    for i, it in enumerate(iterations):
        last = i == len(iterations) - 1
        if it["step_accepted"]:
            if it["lambda"] != lam:
                bad.append(f"iteration {i}: accepted with lambda {it['lambda']}, solved with {lam}")
            if o.has_initial_lambda:
                lam = max(lam * o.lambda_decrease_factor, o.min_lambda)
        else:
            lam = min(lam * o.lambda_increase_factor, o.max_lambda)
            if it["lambda"] != lam:
                bad.append(f"iteration {i}: rejected, reports {it['lambda']}, expected {lam}")
            if lam >= o.max_lambda and not last:
                bad.append(f"iteration {i}: lambda reached max_lambda but the loop went on")
    return bad


def test_lambda_bookkeeping():
    """Reported lambda per iteration, with clamping at both bounds and the
    loop's stops, on scenes that accept, reject on feasibility and fail the
    solve (singular at lambda = 0)."""
    far = scenes.make_problem(scenes.pinhole_camera(), [(10, scenes.pose_at([0, 0, 0]))],
                              [(1, np.array([1.0, 0.0, 0.1]))], [(10, 1, np.array([105320.0, 240.0]))],
                              fixed_poses=(10,))
    o = vba.default_options(max_iterations=10, initial_lambda=1e-4, lambda_increase_factor=10.0, max_lambda=5e-3)
    out = oracle.optimize(far, o)
    assert out.status == abi.OK and not lambda_rule_violations(out.iterations, o)
    assert [it["lambda"] for it in out.iterations] == [min(1e-4 * 10.0, 5e-3), 5e-3]  # stops at the bound
    # singular at lambda = 0 (one observation of a free landmark): the solve
    # fails every time, lambda stays 0 and the loop runs to the budget
    o = vba.default_options(max_iterations=10, has_initial_lambda=0)
    out = oracle.optimize(far, o)
    assert out.status == abi.OK and len(out.iterations) == 10
    assert not lambda_rule_violations(out.iterations, o)
    # a feasibility rejection without an initial lambda stops at once
    two = scenes.make_problem(scenes.pinhole_camera(),
                              [(10, scenes.pose_at([0, 0, 0])), (11, scenes.pose_at([0.3, 0, 0]))],
                              [(1, np.array([1.0, 0.0, 0.1]))],
                              [(10, 1, np.array([105320.0, 240.0])), (11, 1, np.array([-90000.0, 300.0]))],
                              fixed_poses=(10, 11))
    out = oracle.optimize(two, o)
    assert out.status == abi.OK and len(out.iterations) == 1 and not out.iterations[0]["step_accepted"]
    assert out.iterations[0]["cost_after"] != out.iterations[0]["cost_before"]  # a solved, rejected step
    # accepted steps down to the lower bound
    drift = scenes.truth_bundle(drift30=[1.3, 0.3, 0.2])
    o = vba.default_options(max_iterations=20, initial_lambda=1e-4, lambda_decrease_factor=0.5, min_lambda=3e-5)
    out = oracle.optimize(drift, o)
    assert out.status == abi.OK and all(it["step_accepted"] for it in out.iterations)
    assert not lambda_rule_violations(out.iterations, o)
    assert [it["lambda"] for it in out.iterations][:4] == [1e-4, 5e-5, 3e-5, 3e-5]
    # without an initial lambda accepted steps keep lambda at 0
    o = vba.default_options(max_iterations=20, has_initial_lambda=0)
    out = oracle.optimize(drift, o)
    assert out.status == abi.OK and out.iterations and all(it["lambda"] == 0.0 for it in out.iterations)


@pytest.mark.parametrize("name,policy", [("loop_accept", abi.PCG_MAXIT_ACCEPT_ITERATE),
                                         ("loop_reject", abi.PCG_MAXIT_REJECT)])
def test_lambda_bookkeeping_of_the_loop_traces(name, policy):
    """The stored 1,200-camera loop traces (accepts, cost rejections and, under
    the reject-at-cap rule, failed solves) follow the same damping rule."""
    with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "traces", f"{name}.npz"),
                 allow_pickle=False) as z:
        acc, lam = z["accepted"], z["lam"]
    o = workloads.loop_options(policy=policy)
    its = [{"step_accepted": bool(a), "lambda": float(l)} for a, l in zip(acc, lam)]
    assert 0 < int(acc.sum()) < len(acc)
    assert not lambda_rule_violations(its, o)


def test_non_spd_preconditioner_is_rejected():
    """A Schur block that is not positive definite fails the 6x6 Cholesky
    (implicit_schur.rs:279-287, test at :1444)."""
    cross = np.zeros((6, 3))
    cross[0, 0] = cross[1, 1] = cross[2, 2] = 0.8
    ns = oracle.System(diag=[np.eye(6), np.eye(6)], b_p=np.ones(12), h_ll=[0.5 * np.eye(3)],
                       b_l=[[0.0, 0.0, 0.0]], cross=[[(0, cross), (1, cross)]])
    st, *_ = ns.operator(0.0)
    assert st == -4


# ---------------------------------------------------------------------------
# Robust kernels (pose_graph.rs:87-140; bundle_adjustment.rs:1445-1485)
# ---------------------------------------------------------------------------
def test_robust_kernel_values():
    assert oracle.robust(abi.KERNEL_HUBER, 2.0, 0.0)[1] == 1.0
    assert oracle.robust(abi.KERNEL_HUBER, 2.0, 4.0)[1] == 1.0
    assert abs(oracle.robust(abi.KERNEL_HUBER, 2.0, 9.0)[1] - 2.0 / 3.0) < 1e-12
    assert abs(oracle.robust(abi.KERNEL_HUBER, 10.0, 5000.0)[0] - (20.0 * np.sqrt(5000.0) - 100.0)) < 1e-9
    assert abs(oracle.robust(abi.KERNEL_CAUCHY, 10.0, 5000.0)[0] - 100.0 * np.log(51.0)) < 1e-9
    assert abs(oracle.robust(abi.KERNEL_CAUCHY, 1.0, 3.0)[1] - 0.25) < 1e-15
    # Huber weight is the derivative of its cost (visual_jacobian_audit_tests.rs:141)
    for s in [0.5, 3.9, 4.1, 25.0, 400.0]:
        h = 1e-6 * s
        c_plus = oracle.robust(abi.KERNEL_HUBER, 2.0, s + h)[0]
        c_minus = oracle.robust(abi.KERNEL_HUBER, 2.0, s - h)[0]
        assert abs((c_plus - c_minus) / (2 * h) - oracle.robust(abi.KERNEL_HUBER, 2.0, s)[1]) < 1e-6


def test_robust_cost_clips_outlier_to_analytic_huber_value():
    ba = scenes.over_determined_bundle(True)
    squared, _ = oracle.cost(ba)
    huber, _ = oracle.cost(ba, abi.KERNEL_HUBER, 10.0)
    cauchy, _ = oracle.cost(ba, abi.KERNEL_CAUCHY, 10.0)
    assert abs(squared - 5000.0) < 1e-6
    assert huber < squared and cauchy < huber
    assert abs(huber - 1314.0) < 1.0


# ---------------------------------------------------------------------------
# Projection (jacobians.rs:29-124; camera.rs:351-507, 852-941)
# ---------------------------------------------------------------------------
def test_projection_closed_forms():
    pin = vba.camera(abi.CAMERA_PINHOLE, [500.0, 500.0, 320.0, 240.0])
    uv, J = oracle.mono_project(pin, [1.0, -0.5, 5.0])
    assert np.allclose(uv, [420.0, 190.0], atol=1e-12)
    assert np.allclose(J, [[100.0, 0.0, -20.0], [0.0, 100.0, 10.0]], atol=1e-12)
    # a point on the optical axis projects to the principal point for every lens
    for cam in [pin, vba.camera(abi.CAMERA_SIMPLE_RADIAL, [500.0, 320.0, 240.0, 0.3]),
                vba.camera(abi.CAMERA_OPENCV, [510.0, 490.0, 321.0, 239.0, -0.1, 0.05, 0.01, -0.02])]:
        uv = oracle.camera_project(cam, [0.0, 0.0, 3.0])
        assert np.allclose(uv, [cam.params[1] if cam.model == abi.CAMERA_SIMPLE_RADIAL else cam.params[2],
                                cam.params[2] if cam.model == abi.CAMERA_SIMPLE_RADIAL else cam.params[3]],
                           atol=1e-12)
    # SIMPLE_RADIAL [f, cx, cy, k]: x = 0.5, r^2 = 0.25, d = 1 + 0.01 * 0.25
    rad = vba.camera(abi.CAMERA_SIMPLE_RADIAL, [500.0, 320.0, 240.0, 0.01])
    uv = oracle.camera_project(rad, [0.5, 0.0, 1.0])
    assert abs(uv[0] - 570.625) < 1e-12 and abs(uv[1] - 240.0) < 1e-12
    # behind the camera: None
    for cam in [pin, rad]:
        assert oracle.camera_project(cam, [0.1, 0.1, -1.0]) is None
        assert oracle.mono_project(cam, [0.1, 0.1, -1.0]) is None


def test_opencv_projection_known_answer():
    """imported_opencv_camera_applies_tangential_distortion
    (crates/io/tests/colmap.rs:651-668): the pixel OpenCV 4.10
    cv2.projectPoints gives for a real COLMAP OPENCV camera, to 1e-8 px, with
    the tangential terms (a k1/k2-only projection misses it by pixels)."""
    cam = vba.camera(abi.CAMERA_OPENCV, [1353.09, 1338.03, 962.7, 539.93, 0.0346, -0.0235, 0.0012, -0.0008])
    uv = oracle.camera_project(cam, [0.5, -0.3, 1.0])
    assert abs(uv[0] - 1643.9695843529998) < 1e-8 and abs(uv[1] - 136.0453501334) < 1e-8, uv
    uv2, _ = oracle.mono_project(cam, [0.5, -0.3, 1.0])
    assert abs(uv2[0] - 1643.9695843529998) < 1e-8 and abs(uv2[1] - 136.0453501334) < 1e-8, uv2
    radial_only = vba.camera(abi.CAMERA_OPENCV, [1353.09, 1338.03, 962.7, 539.93, 0.0346, -0.0235, 0.0, 0.0])
    assert np.abs(oracle.camera_project(radial_only, [0.5, -0.3, 1.0]) - uv).max() > 1.0


LENSES = [
    vba.camera(abi.CAMERA_PINHOLE, [512.5, 498.25, 321.0, 239.5]),
    vba.camera(abi.CAMERA_SIMPLE_RADIAL, [500.0, 320.0, 240.0, 0.01]),
    vba.camera(abi.CAMERA_PINHOLE, [512.5, 498.25, 321.0, 239.5, -0.17, 0.06]),
    vba.camera(abi.CAMERA_OPENCV, [500.0, 500.0, 320.0, 240.0, -0.15, 0.05, 0.004, -0.003]),
    vba.camera(abi.CAMERA_OPENCV, [500.0, 500.0, 320.0, 240.0, -0.15, 0.05, 0.0, 0.0]),
]


def sample_points():
    """mono_projection_tests.rs:15-26."""
    return [np.array([-1.3 + 0.41 * i, -0.9 + 0.43 * j, 1.5 + 0.37 * (i + j)]) for i in range(7) for j in range(5)]


@pytest.mark.parametrize("cam", LENSES, ids=lambda c: f"model{c.model}")
def test_projection_jacobian_matches_finite_differences(cam):
    """mono_projection_tests.rs:126-191."""
    for xc in sample_points():
        uv, J = oracle.mono_project(cam, xc)
        uv_cost = oracle.camera_project(cam, xc)
        assert np.allclose(uv, uv_cost, rtol=0, atol=1e-10)
        h = 1e-6
        for axis in range(3):
            p, m = xc.copy(), xc.copy()
            p[axis] += h
            m[axis] -= h
            num = (oracle.camera_project(cam, p) - oracle.camera_project(cam, m)) / (2 * h)
            for row in range(2):
                assert abs(J[row, axis] - num[row]) < 1e-5 * (1 + abs(J[row, axis]))


def test_radial_projection_equals_cost_projection_bitwise():
    """The cost evaluates Camera::project: identical bits for the radial
    closed form (mono_projection_tests.rs:118-120)."""
    for cam in LENSES[1:3] + LENSES[4:5]:
        for xc in sample_points():
            uv, _ = oracle.mono_project(cam, xc)
            assert np.array_equal(uv, oracle.camera_project(cam, xc))


# ---------------------------------------------------------------------------
# SE3 (se3.rs:57-134, tests :153-211)
# ---------------------------------------------------------------------------
def test_se3_exp_closed_forms_and_log_round_trip():
    # pure translation
    p = oracle.se3_exp([0.3, -0.2, 0.1, 0.0, 0.0, 0.0])
    assert np.allclose(p, [1, 0, 0, 0, 0.3, -0.2, 0.1], atol=1e-15)
    # rotation about z by theta: q = (cos th/2, 0, 0, sin th/2)
    th = 0.7
    p = oracle.se3_exp([0, 0, 0, 0, 0, th])
    assert np.allclose(p[:4], [np.cos(th / 2), 0, 0, np.sin(th / 2)], atol=1e-15)
    assert np.allclose(scenes.quat_to_matrix(p[:4]), scenes.rodrigues(np.array([0, 0, th])), atol=1e-15)
    for xi in [[0.1, -0.2, 0.3, 0.01, 0.02, -0.03], [1.0, 2.0, -0.5, 0.5, -1.2, 0.4], [0, 0, 0, 1e-9, 0, 0],
               [0.2, 0.0, 0.0, 2.5, 0.3, -0.1]]:
        back = oracle.se3_log(oracle.se3_exp(xi))
        assert np.allclose(back, xi, atol=1e-12)
    # exp(xi) applied to a point equals the 4x4 matrix exponential
    xi = np.array([0.4, -0.1, 0.2, 0.3, -0.2, 0.5])
    Xi = np.zeros((4, 4))
    Xi[:3, :3] = [[0, -xi[5], xi[4]], [xi[5], 0, -xi[3]], [-xi[4], xi[3], 0]]
    Xi[:3, 3] = xi[:3]
    from scipy.linalg import expm

    T = expm(Xi)
    p = oracle.se3_exp(xi)
    assert np.allclose(scenes.quat_to_matrix(p[:4]), T[:3, :3], atol=1e-14)
    assert np.allclose(p[4:], T[:3, 3], atol=1e-14)


# ---------------------------------------------------------------------------
# Per-observation Jacobians vs finite differences of the residual
# (visual_jacobian_audit_tests.rs:49; right perturbation T <- T Exp(xi)).
# ---------------------------------------------------------------------------
def residual(cam, pose, X, uv):
    return oracle.camera_project(cam, scenes.transform(pose, X)) - uv


@pytest.mark.parametrize("cam", LENSES, ids=lambda c: f"model{c.model}")
def test_linearization_matches_finite_differences(cam):
    rng = np.random.default_rng(7)
    poses = [(1, scenes.pose_from_rt(scenes.rodrigues(rng.normal(0, 0.2, 3)), rng.normal(0, 0.3, 3)))]
    lms = [(i + 1, np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(4, 6)])) for i in range(8)]
    obs = [(1, l, np.array([300.0, 250.0]) + rng.normal(0, 3, 2)) for l, _ in lms]
    cam.width, cam.height = 640, 480
    ba = scenes.make_problem(cam, poses, lms, obs)
    r, jp, jl, w, valid = oracle.linearize(ba)
    assert valid.all() and np.all(w == 1.0)
    h = 1e-6
    for k, (lid, X) in enumerate(lms):
        pose = poses[0][1]
        assert np.allclose(r[k], residual(cam, pose, X, obs[k][2]), atol=1e-9)
        for a in range(6):
            xi = np.zeros(6)
            xi[a] = h
            rp = residual(cam, oracle.se3_compose(pose, oracle.se3_exp(xi)), X, obs[k][2])
            xi[a] = -h
            rm = residual(cam, oracle.se3_compose(pose, oracle.se3_exp(xi)), X, obs[k][2])
            assert np.allclose(jp[k][:, a], (rp - rm) / (2 * h), rtol=1e-5, atol=1e-4)
        for a in range(3):
            d = np.zeros(3)
            d[a] = h
            num = (residual(cam, pose, X + d, obs[k][2]) - residual(cam, pose, X - d, obs[k][2])) / (2 * h)
            assert np.allclose(jl[k][:, a], num, rtol=1e-5, atol=1e-4)


# ---------------------------------------------------------------------------
# End-to-end LM (pipelines/slam/tests/bundle_adjustment.rs)
# ---------------------------------------------------------------------------
def test_cost_is_zero_at_truth():
    c, nonproj = oracle.cost(scenes.truth_bundle())
    assert c < 1e-20 and nonproj == 0


def test_drifted_pose_recovered():
    """bundle_adjustment.rs:367-402."""
    ba = scenes.truth_bundle(drift30=[1.05, 0.02, 0.03])
    assert oracle.cost(ba)[0] > 1e-3
    out = oracle.optimize(ba, vba.default_options())
    assert out.status == abi.OK
    assert out.result.final_cost < 1e-12 and out.result.final_cost < out.result.initial_cost
    assert out.result.converged
    centre = vba.camera_centres(out.poses)[2]
    assert np.linalg.norm(centre - [1.0, 0.0, 0.0]) < 1e-6


def test_feasibility_gate_rejects_step_behind_camera():
    """bundle_adjustment.rs:1004-1032."""
    ba = scenes.make_problem(scenes.pinhole_camera(), [(10, scenes.pose_at([0, 0, 0]))],
                             [(1, np.array([1.0, 0.0, 0.1]))], [(10, 1, np.array([105320.0, 240.0]))],
                             fixed_poses=(10,))
    out = oracle.optimize(ba, vba.default_options(max_iterations=1))
    assert out.status == abi.OK
    assert np.array_equal(out.landmarks[0], [1.0, 0.0, 0.1])
    assert len(out.iterations) == 1 and not out.iterations[0]["step_accepted"]
    assert out.result.final_cost == out.result.initial_cost


def test_dense_and_matrix_free_agree_on_small_scene():
    """bundle_adjustment.rs:1160-1200 (sparse == dense), here dense ==
    matrix-free with a tight PCG."""
    ba = scenes.truth_bundle(drift30=[1.04, 0.01, -0.02])
    d = oracle.optimize(ba, vba.default_options())
    m = oracle.optimize(ba, vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE))
    assert d.status == m.status == abi.OK
    assert d.result.converged == m.result.converged
    assert abs(d.result.final_cost - m.result.final_cost) < 1e-12
    assert np.linalg.norm(vba.camera_centres(d.poses)[2] - vba.camera_centres(m.poses)[2]) < 1e-9


def test_first_lm_step_equals_brute_force_full_solve():
    """Schur reduction + back-substitution + update against the full damped
    normal system solved by numpy from the oracle's per-observation
    linearization (itself pinned by finite differences above)."""
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=5, num_landmarks=40, seed=11)
    ba = s.problem
    lam = 1e-4
    r, jp, jl, w, valid = oracle.linearize(ba)
    P = len(ba.pose_ids)
    var_pose = {i: k for k, i in enumerate([i for i in range(P) if not ba.pose_fixed[i]])}
    Pv, L = len(var_pose), len(ba.landmark_ids)
    pose_index = {int(pid): i for i, pid in enumerate(ba.pose_ids)}
    lm_index = {int(lid): i for i, lid in enumerate(ba.landmark_ids)}
    n = 6 * Pv + 3 * L
    H = np.zeros((n, n))
    g = np.zeros(n)
    for k in range(ba.num_observations):
        if not valid[k]:
            continue
        p = var_pose.get(pose_index[int(ba.obs_pose_ids[k])])
        l = lm_index[int(ba.obs_landmark_ids[k])]
        J = np.zeros((2, n))
        if p is not None:
            J[:, 6 * p:6 * p + 6] = jp[k]
        J[:, 6 * Pv + 3 * l:6 * Pv + 3 * l + 3] = jl[k]
        H += w[k] * J.T @ J
        g += w[k] * J.T @ r[k]
    delta = np.linalg.solve(H + lam * np.eye(n), -g)
    out = oracle.optimize(ba, vba.default_options(max_iterations=1))
    assert out.iterations[0]["step_accepted"]
    for i, k in var_pose.items():
        expect = oracle.se3_compose(ba.poses[i], oracle.se3_exp(delta[6 * k:6 * k + 6]))
        assert np.allclose(out.poses[i], expect, atol=1e-9)
    assert np.allclose(out.landmarks, ba.landmarks + delta[6 * Pv:].reshape(-1, 3), atol=1e-9)


@pytest.mark.parametrize("case", ["loop_dense", "loop_pcg_accept", "loop_pcg_reject", "circle_dense"])
def test_parallel_oracle_is_bitwise_equal_to_serial(case):
    """parallel_ba_tests.rs:113-160: the threaded solve equals the serial one
    bit for bit. The problems exceed 65,536 observations so that the threaded
    assembly, cost, Schur reduction, implicit operator, preconditioner and
    back-substitution all run."""
    if case == "circle_dense":
        s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=12, num_landmarks=6000, seed=5)
        o = vba.default_options(max_iterations=4)
    else:
        s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=90, num_landmarks=10000, seed=5, scene_scale=30.0)
        if case == "loop_dense":
            o = vba.default_options(max_iterations=4)
        else:
            o = vba.default_options(max_iterations=4, linear_solver=abi.SOLVER_MATRIX_FREE, max_pcg_iterations=25,
                                    pcg_relative_tolerance=1e-6, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                                    pcg_maxit_policy=(abi.PCG_MAXIT_ACCEPT_ITERATE if case.endswith("accept")
                                                      else abi.PCG_MAXIT_REJECT))
    assert s.problem.num_observations > 65536
    a = oracle.optimize(s.problem, o, threads=1)
    b = oracle.optimize(s.problem, o, threads=5)
    assert a.status == b.status == abi.OK
    assert a.result.final_cost == b.result.final_cost
    assert np.array_equal(a.poses, b.poses) and np.array_equal(a.landmarks, b.landmarks)
    assert a.iterations == b.iterations
    if case != "circle_dense" and case != "loop_dense":
        assert any(it["pcg_iterations"] > 0 for it in a.iterations)


def test_kernel_none_matches_gauss_newton_weights():
    """bundle_adjustment.rs:1371: RobustKernel::None gives unit weights."""
    ba = scenes.over_determined_bundle(True)
    _, _, _, w, valid = oracle.linearize(ba)
    assert valid.all() and np.all(w == 1.0)
    _, _, _, wh, _ = oracle.linearize(ba, abi.KERNEL_HUBER, 10.0)
    assert np.sum(wh < 1.0) == 1


def test_error_codes():
    cam = scenes.pinhole_camera()
    empty = scenes.make_problem(cam, [], [], [])
    assert oracle.optimize(empty, vba.default_options()).status == abi.ERR_NO_POSES
    one = scenes.make_problem(cam, [(1, scenes.pose_at([0, 0, 0]))], [], [])
    assert oracle.optimize(one, vba.default_options()).status == abi.ERR_NO_OBSERVATIONS
    nolm = scenes.make_problem(cam, [(1, scenes.pose_at([0, 0, 0]))], [], [(1, 99, np.array([320.0, 240.0]))])
    assert oracle.optimize(nolm, vba.default_options()).status == abi.ERR_NO_LANDMARKS
    ba = scenes.truth_bundle()
    bad = ba.copy()
    bad.obs_pose_ids[3] = 77
    out = oracle.optimize(bad, vba.default_options())
    assert out.status == abi.ERR_MISSING_POSE and out.result.error_id == 77
    bad = ba.copy()
    bad.obs_landmark_ids[5] = 1234
    out = oracle.optimize(bad, vba.default_options())
    assert out.status == abi.ERR_MISSING_LANDMARK and out.result.error_id == 1234
    allfixed = ba.copy()
    allfixed.pose_fixed[:] = 1
    allfixed.landmark_fixed[:] = 1
    assert oracle.optimize(allfixed, vba.default_options()).status == abi.ERR_ALL_POSES_FIXED


def test_matrix_free_entry_validation():
    """optimize.rs:382-626."""
    ba = scenes.truth_bundle(fixed=(), drift30=[1.05, 0.0, 0.0])
    mf = vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE)
    assert oracle.optimize(ba, mf).status == abi.ERR_INELIGIBLE  # no gauge anchor
    ba = scenes.truth_bundle(drift30=[1.05, 0.0, 0.0])
    assert oracle.optimize(ba, vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE,
                                                   has_initial_lambda=0)).status == abi.ERR_INVALID_CONFIGURATION
    assert oracle.optimize(ba, vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE,
                                                   max_pcg_iterations=0)).status == abi.ERR_INVALID_CONFIGURATION
    assert oracle.optimize(ba, vba.default_options(linear_solver=abi.SOLVER_MATRIX_FREE,
                                                   robust_kernel=abi.KERNEL_HUBER,
                                                   robust_scale=0.0)).status == abi.ERR_INVALID_CONFIGURATION


def test_circle_lm_converges_and_recovers_cameras():
    """A reduced CIRCLE workload: LM converges and the camera centres move
    towards the truth."""
    s = vba.synthesize(abi.SYNTH_CIRCLE, num_cameras=10, num_landmarks=200, seed=1)
    out = oracle.optimize(s.problem, vba.default_options())
    assert out.status == abi.OK and out.result.final_cost < 0.05 * out.result.initial_cost
    before = np.linalg.norm(vba.camera_centres(s.problem.poses) - vba.camera_centres(s.truth_poses), axis=1)
    after = np.linalg.norm(vba.camera_centres(out.poses) - vba.camera_centres(s.truth_poses), axis=1)
    assert after.max() < before.max()


def test_solve_normal_equations_cholesky_lu_and_singular():
    """solve_normal_equations (pose_graph.rs:2259-2270): Cholesky on the
    lower triangle when H is positive definite, LU with partial pivoting when
    it is not, None when LU meets a zero pivot."""
    rng = np.random.default_rng(11)
    a = rng.standard_normal((30, 30))
    h = a @ a.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    x = oracle.solve_normal_equations(h, b)
    assert np.abs(x - np.linalg.solve(h, b)).max() <= 1e-12 * np.abs(x).max()
    # Cholesky reads the lower triangle only
    hu = h.copy()
    hu[np.triu_indices(30, 1)] = 1e3
    assert np.array_equal(oracle.solve_normal_equations(hu, b), x)
    # a zero first pivot: Cholesky fails, LU swaps rows; exact answer
    h3 = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
    assert np.array_equal(oracle.solve_normal_equations(h3, np.array([1.0, 2.0, 3.0])), [2.0, 1.0, -3.0])
    # symmetric indefinite: LU against numpy (LAPACK getrf, the same pivots)
    q, _ = np.linalg.qr(rng.standard_normal((25, 25)))
    hi = (q * np.linspace(-2.0, 2.5, 25)) @ q.T
    bi = rng.standard_normal(25)
    xi = oracle.solve_normal_equations(hi, bi)
    assert np.abs(xi - np.linalg.solve(hi, bi)).max() <= 1e-9 * np.abs(xi).max()
    # singular: a zero row and column
    hs = h.copy()
    hs[4, :] = 0.0
    hs[:, 4] = 0.0
    assert oracle.solve_normal_equations(hs, b) is None


def column_sweep_cholesky(a):
    """nalgebra's Cholesky::new as a plain column sweep, in Python floats
    (IEEE doubles, every product and sum rounded on its own): column j gets
    -L(j, k) L(i, k) added for k ascending, then its diagonal's square root
    divides it."""
    n = len(a)
    m = [list(map(float, row)) for row in a]
    for j in range(n):
        for k in range(j):
            f = -m[j][k]
            for i in range(j, n):
                m[i][j] = f * m[i][k] + m[i][j]
        d = m[j][j]
        if not (d != 0.0 and d >= 0.0):
            return None
        r = math.sqrt(d)
        m[j][j] = r
        for i in range(j + 1, n):
            m[i][j] = m[i][j] / r
    return np.tril(np.array(m))


@pytest.mark.parametrize("n,threads", [(5, 1), (63, 1), (64, 1), (65, 1), (150, 1), (300, 4)])
def test_blocked_cholesky_is_the_column_sweep(n, threads):
    """The oracle's blocked, threaded Cholesky does each entry's operations
    in the column sweep's order: the factors are bitwise equal."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    h = a @ a.T + n * np.eye(n)
    ref = column_sweep_cholesky(h)
    got = oracle.cholesky(h, threads)
    assert np.array_equal(got, ref)
    bad = h.copy()
    bad[n // 2, n // 2] = -1.0
    assert oracle.cholesky(bad, threads) is None
