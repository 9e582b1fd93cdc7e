"""bench.py --replay (crates/ba-gpu/examples/ba_gpu_replay.rs): a problem
dump written by the library, read back and windowed, solved on the device and
by the CPU oracle with the same options. The CPU test pins the reported RMS
reprojection error to the oracle's cost with no robust kernel (the
reference's robust_cost(None), problem.rs:497-526)."""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import vba
from vba import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def write_dump(path, cameras=40, landmarks=2500, seed=5):
    s = vba.synthesize(abi.SYNTH_LOOP, num_cameras=cameras, num_landmarks=landmarks, seed=seed,
                       scene_scale=0.2 * cameras)
    with vba.Model.from_problem(s.problem) as m:
        m.write_ba_problem(path)
    return s.problem


def test_rms_is_the_unrobust_cost_per_observation(tmp_path):
    path = tmp_path / "problem.txt"
    write_dump(path)
    with vba.Model.read_ba_problem(path) as m:
        m.keep_local_window(12)
        problem, weights = m.problem()
    assert weights is None and 0 < np.count_nonzero(problem.pose_fixed) < len(problem.pose_ids)
    rng = np.random.default_rng(1)
    lms = problem.landmarks + rng.normal(scale=0.05, size=problem.landmarks.shape)
    moved = problem.with_state(problem.poses, lms)
    cost, nonproj = oracle.cost(moved)
    rms = bench.rms_reprojection(problem, problem.poses, lms)
    assert nonproj == 0
    assert abs(rms - math.sqrt(cost / problem.num_observations)) <= 1e-12 * rms


@pytest.mark.gpu
def test_replay_matches_the_oracle(tmp_path):
    path = tmp_path / "problem.txt"
    write_dump(path)
    for local, solver in ((12, "dense"), (0, "sparse"), (0, "matrix-free")):
        cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--replay", str(path), "--iterations", "8",
               "--replay-solver", solver] + (["--local", str(local)] if local else [])
        print("replay", local, solver, flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
        rec = json.loads(r.stdout.strip().splitlines()[-1])
        print(r.stderr.strip().splitlines()[-2:], flush=True)
        assert rec["device"]["status"] == abi.OK and rec["cpu"]["status"] == abi.OK
        assert rec["same_accept_vector"], (rec["device"]["accept_vector"], rec["cpu"]["accept_vector"])
        assert rec["rel_cost_vs_cpu"] <= 1e-6
        assert abs(rec["device"]["rms_px"] - rec["cpu"]["rms_px"]) <= 1e-6 * rec["cpu"]["rms_px"]
