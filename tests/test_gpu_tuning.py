"""The PCG launch schedule (vba_tuning) and its autotuner. The device decides
when PCG stops and launches past the stop do nothing, so every schedule must
give the default schedule's result bit for bit: iteration records, final
cost and final state."""
import numpy as np
import pytest

import vba
from vba import abi

pytestmark = pytest.mark.gpu

# (chunk, lookahead, predict, persistent product waves per CU); the persistent
# product at several residencies, including more waves than resident
SCHEDULES = [(10, 0, False, 0), (1, 0, False, 0), (3, 2, True, 0), (7, 1, False, 0), (25, 0, True, 0),
             (4, 2, False, 0), (10, 0, False, 16), (3, 2, True, 4), (7, 1, False, 1), (20, 1, True, 32)]


def loop_options(solver, precision, max_pcg=100):
    return vba.default_options(linear_solver=solver, max_iterations=6, max_pcg_iterations=max_pcg,
                               pcg_relative_tolerance=1e-4, robust_kernel=abi.KERNEL_HUBER, robust_scale=1.0,
                               pcg_maxit_policy=abi.PCG_MAXIT_ACCEPT_ITERATE, precision=precision)


def solve_all(device, pp, o):
    outs = []
    for chunk, ahead, predict, waves in SCHEDULES:
        device.set_tuning(chunk, ahead, predict, waves)
        out = pp.solve(o)
        assert out.status == abi.OK, out.message
        outs.append((out, pp.state()))
    device.set_tuning()
    return outs


@pytest.mark.parametrize("solver", [abi.SOLVER_MATRIX_FREE, abi.SOLVER_PCG_EXPLICIT])
@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
@pytest.mark.parametrize("max_pcg", [100, 7])
def test_schedules_give_identical_results(device, solver, precision, max_pcg):
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=80, num_landmarks=6000, seed=2, scene_scale=16.0)
    o = loop_options(solver, precision, max_pcg)
    pp = device.prepare(s.problem, o)
    outs = solve_all(device, pp, o)
    (ref, (p0, l0)) = outs[0]
    assert ref.result.total_pcg_iterations > 0
    for out, (p, l) in outs[1:]:
        assert out.iterations == ref.iterations
        assert out.result.final_cost == ref.result.final_cost
        assert out.result.total_pcg_iterations == ref.result.total_pcg_iterations
        np.testing.assert_array_equal(p, p0)
        np.testing.assert_array_equal(l, l0)
    pp.free()


def test_autotune_keeps_the_result(device):
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=80, num_landmarks=6000, seed=5, scene_scale=16.0)
    o = loop_options(abi.SOLVER_MATRIX_FREE, abi.PRECISION_MIXED)
    pp = device.prepare(s.problem, o)
    device.set_tuning()
    ref = pp.solve(o)
    p0, l0 = pp.state()
    tuned = pp.autotune(o, lm_iterations=2)
    assert tuned["seconds"] <= tuned["default_seconds"]
    assert device.tuning() == tuned["tuning"]
    out = pp.solve(o)
    assert out.iterations == ref.iterations and out.result.final_cost == ref.result.final_cost
    p, l = pp.state()
    np.testing.assert_array_equal(p, p0)
    np.testing.assert_array_equal(l, l0)
    device.set_tuning()
    pp.free()


def test_tuning_rejects_out_of_range(device):
    with pytest.raises(ValueError):
        device.set_tuning(0, 0, False)
    with pytest.raises(ValueError):
        device.set_tuning(10, 3, False)
    with pytest.raises(ValueError):
        device.set_tuning(10, 0, False, 33)
    assert device.tuning() == {"pcg_chunk": 10, "pcg_lookahead": 0, "pcg_predict": 0, "apply_waves_per_cu": 0}


def test_prepared_solves_decline_out_of_range_options(device):
    """A prepared problem or batch applies the option checks of vba_check to
    every solve: out-of-range enums and intrinsics refinement decline
    instead of running with a default."""
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=20, num_landmarks=600, seed=5, scene_scale=12.0)
    o = loop_options(abi.SOLVER_MATRIX_FREE, abi.PRECISION_FP64)
    pp = device.prepare(s.problem, o)
    bad = [dict(linear_solver=4), dict(linear_solver=9), dict(robust_kernel=5), dict(precision=3),
           dict(pcg_maxit_policy=4), dict(refine_intrinsics=1)]
    for fields in bad:
        ob = loop_options(abi.SOLVER_MATRIX_FREE, abi.PRECISION_FP64)
        for k, v in fields.items():
            setattr(ob, k, v)
        out = pp.solve(ob)
        assert out.status == abi.DECLINED, fields
        assert "declined" in out.message
        with pytest.raises(RuntimeError):
            pp.autotune(ob, lm_iterations=1)
    assert pp.solve(o).status == abi.OK
    w = vba.synthesize(abi.SYNTH_STEREO_WINDOW, num_cameras=6, num_landmarks=300, seed=3).problem
    od = vba.default_options(max_iterations=3)
    pb = device.prepare_batch([w], od)
    ob = vba.default_options(max_iterations=3, robust_kernel=5)
    assert pb.solve(ob).status == abi.DECLINED
    assert pb.solve(od).status == abi.OK


@pytest.mark.parametrize("stride", [1, 3])
def test_profiling_keeps_the_result_and_times_each_family(device, stride):
    """Per-kernel timing (vba_profile_enable): the PCG loop's product, pose
    and step launches carry their events on the dispatch, the other kernels
    are bracketed on the stream. Timing changes no result, and each family
    reports sampled launches with positive, plausible times; no more product
    launches than PCG iterations run (launches past the device's stop are
    not counted)."""
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=80, num_landmarks=6000, seed=3, scene_scale=16.0)
    o = loop_options(abi.SOLVER_MATRIX_FREE, abi.PRECISION_MIXED)
    pp = device.prepare(s.problem, o)
    device.set_tuning()
    device.profile(False)
    ref = pp.solve(o)
    assert ref.status == abi.OK, ref.message
    p0, l0 = pp.state()
    device.profile(True, stride=stride)
    out = pp.solve(o)
    device.profile(False, reset=False)
    assert out.status == abi.OK, out.message
    p1, l1 = pp.state()
    assert out.iterations == ref.iterations
    assert out.result.final_cost == ref.result.final_cost
    np.testing.assert_array_equal(p1, p0)
    np.testing.assert_array_equal(l1, l0)
    pcg = out.result.total_pcg_iterations
    assert pcg > 0
    for name in ("schur_apply", "pcg_pose", "pcg_step", "linearize", "cost"):
        q = device.profile_query(name)
        assert q["launches"] > 0, name
        assert q["seconds"] > 0.0, name
        assert 1e-7 < q["seconds"] / q["launches"] < 1e-2, (name, q)
    assert device.profile_query("schur_apply")["launches"] <= pcg
    pp.free()
