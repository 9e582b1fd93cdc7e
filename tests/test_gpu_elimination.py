"""The device point elimination, vba_reduced_eliminate
(ReducedSolveAccelerator::eliminate, pipelines/slam/src/ba_accel.rs:71-157),
against the oracle's restatement of the CPU rig solver's eliminate_by_output
(oracle_reduced_eliminate; rig_ba_solver.rs:2551-2660), which
tests/test_elimination_oracle.py pins to the dense Schur complement.

The device repeats the CPU's arithmetic (the product hcp V^-1 with rounded
multiplies and adds, fused multiply-add chains per point in ascending point
order, the shard partial sums), so the run blocks and aggregate sums are
bitwise equal to the oracle's; the FP32 blocks it keeps for the solve are
then the FP32 rounding of the oracle's blocks, and a solve over them equals
a solve handed those blocks explicitly."""
import numpy as np
import pytest

import oracle
import reduced_pcg as R
import scenes

pytestmark = pytest.mark.gpu


def diag_and_rhs(el, blocks, seed=1):
    """A diagonally dominant damped diagonal for the eliminated blocks."""
    rng = np.random.default_rng(seed)
    n = el.num_nodes
    row = np.zeros(n)
    for (a, b), blk in zip(el.edges.tolist(), blocks):
        w = np.abs(blk).sum()
        row[a] += w
        row[b] += w
    g = rng.standard_normal((n, 6, 6)) * 0.1
    diag = g @ np.swapaxes(g, 1, 2) + np.eye(6)[None] * (row + 1.0)[:, None, None]
    return diag, rng.standard_normal(6 * n)


@pytest.mark.parametrize("seed,no_inverse", [(11, 0.1), (12, 0.0)])
def test_eliminate_equals_the_oracle(device, seed, no_inverse):
    el = scenes.elimination_problem(n=150, n_points=4000, seed=seed, no_inverse=no_inverse)
    blocks, runs_ref, coarse_ref = oracle.reduced_eliminate(el)
    out = device.reduced_eliminate(el)
    assert out is not None, device.last_error()
    runs, coarse = out
    assert np.array_equal(runs, runs_ref), np.abs(runs - runs_ref).max() / np.abs(runs_ref).max()
    assert np.array_equal(coarse, coarse_ref), np.abs(coarse - coarse_ref).max() / np.abs(coarse_ref).max()
    # the solve over the device's blocks equals one handed the oracle's blocks
    diag, rhs = diag_and_rhs(el, blocks)
    got = device.reduced_solve(diag, el.edges, None, rhs, 1e-10, structure_id=el.structure_id)
    assert got is not None, device.last_error()
    ref = device.reduced_solve(diag, el.edges, blocks.astype(np.float32), rhs, 1e-10,
                               structure_id=el.structure_id + 1)
    assert got[1] == ref[1]
    assert np.array_equal(got[0], ref[0])
    x, it = R.pcg(diag, el.edges, blocks.astype(np.float32).astype(np.float64), rhs, 1e-10, 500)
    assert it == got[1]
    assert np.linalg.norm(got[0] - x) <= 1e-9 * np.linalg.norm(x)


def test_new_linearization_of_the_same_structure(device):
    el = scenes.elimination_problem(n=90, n_points=1500, seed=13)
    assert device.reduced_eliminate(el) is not None
    rng = np.random.default_rng(5)
    el.hcp = el.hcp * (1.0 + 0.1 * rng.standard_normal(el.hcp.shape))
    el.has_inverse = el.has_inverse.copy()
    el.has_inverse[::7] = 0
    el.linearization_id = 2
    _, runs_ref, coarse_ref = oracle.reduced_eliminate(el)
    runs, coarse = device.reduced_eliminate(el)
    assert np.array_equal(runs, runs_ref), np.abs(runs - runs_ref).max()
    assert np.array_equal(coarse, coarse_ref), np.abs(coarse - coarse_ref).max()


def test_declines(device):
    el = scenes.elimination_problem(n=60, n_points=500, seed=14)
    blocks, _, _ = oracle.reduced_eliminate(el)
    diag, rhs = diag_and_rhs(el, blocks)
    # no elimination of this structure yet: a solve without blocks is declined
    assert device.reduced_solve(diag, el.edges, None, rhs, 1e-8, structure_id=777) is None
    assert "elimination" in device.last_error()
    # a pair of a point's nodes that is not an edge
    bad = scenes.elimination_problem(n=60, n_points=500, seed=14)
    keep = np.ones(len(bad.edges), bool)
    keep[len(keep) // 2] = False
    bad.edges = bad.edges[keep]
    bad.row_off = np.searchsorted(bad.edges[:, 0], np.arange(bad.num_nodes + 1)).astype(np.uint64)
    bad.run_edges = np.zeros(0, np.uint32)
    bad.extra_edge = np.zeros(0, np.uint32)
    bad.extra_blocks = np.zeros((0, 6, 6))
    bad.structure_id = 4242
    assert device.reduced_eliminate(bad) is None
    assert "not an edge" in device.last_error()
    # structure id 0 cannot name the blocks for a later solve
    el.structure_id = 0
    assert device.reduced_eliminate(el) is None
