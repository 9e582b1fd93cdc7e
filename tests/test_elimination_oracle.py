"""The oracle's point elimination (oracle_reduced_eliminate, restating the
CPU rig solver's eliminate_by_output, rig_ba_solver.rs:2551-2660, and what
ReducedSolveAccelerator::eliminate returns, ba_accel.rs:71-157) pinned
against the dense Schur complement -Hcp Vinv Hcp^T computed by numpy from
the same blocks, which shares none of its loops, orders or index
structures."""
import numpy as np
import pytest

import oracle
import scenes

REL = 1e-12


@pytest.mark.parametrize("seed,no_inverse", [(3, 0.1), (4, 0.0), (5, 0.5)])
def test_eliminated_blocks_equal_the_dense_schur_complement(seed, no_inverse):
    el = scenes.elimination_problem(n=60, n_points=700, seed=seed, no_inverse=no_inverse)
    blocks, runs, coarse = oracle.reduced_eliminate(el)
    ref = scenes.elimination_brute_force(el)
    assert blocks.shape == ref.shape == (len(el.edges), 6, 6)
    assert np.abs(blocks - ref).max() <= REL * np.abs(ref).max()
    # pairs sharing no point carry only their direct coupling (or nothing)
    assert np.count_nonzero(np.abs(blocks).sum(axis=(1, 2))) > 0.9 * len(el.edges)


def test_run_blocks_and_coarse_sums():
    el = scenes.elimination_problem(n=70, n_points=900, seed=6)
    blocks, runs, coarse = oracle.reduced_eliminate(el)
    assert len(el.run_edges) > 0
    assert np.array_equal(runs, blocks[el.run_edges])
    group, nodes = el.coarse
    m = -(-nodes // group)
    assert coarse.shape == (m, m, 6, 6)
    expect = np.zeros((m, m, 6, 6))
    for (a, b), blk in zip(el.edges.tolist(), blocks):
        if a < nodes and b < nodes:
            expect[a // group, b // group] += blk
    for g in range(m):
        for h in range(m):
            if g <= h:
                assert np.abs(coarse[g, h] - expect[g, h]).max() <= REL * max(1.0, np.abs(expect[g, h]).max())
            else:
                assert not coarse[g, h].any()


def test_a_point_without_inverse_contributes_nothing():
    el = scenes.elimination_problem(n=40, n_points=300, seed=7, no_inverse=0.0)
    base, _, _ = oracle.reduced_eliminate(el)
    el.has_inverse = el.has_inverse.copy()
    el.has_inverse[17] = 0
    less, _, _ = oracle.reduced_eliminate(el)
    nodes = el.point_nodes[el.point_off[17]:el.point_off[18]].tolist()
    touched = {(a, b) for i, a in enumerate(nodes) for b in nodes[i + 1:]}
    for e, (s, t) in enumerate(el.edges.tolist()):
        if (s, t) in touched:
            assert not np.array_equal(base[e], less[e])
        else:
            assert np.array_equal(base[e], less[e])
