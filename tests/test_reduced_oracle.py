"""The reduced-camera PCG oracle (oracle/reduced_pcg.py) against the
reference's own checks of its CPU solver (colmap_incremental/
rig_ba_solver.rs iterative_bench, :5202-5320) and against dense solves."""
import numpy as np
import pytest

import reduced_pcg as R


def _rhs(n):
    return (np.arange(6 * n) * 13 % 17).astype(np.float64) - 8.0


def test_cluster_preconditioned_pcg_matches_cholesky():
    """rig_ba_solver.rs:5202-5250: block-Jacobi and runs-of-16 PCG both reach
    the dense solution, the runs in fewer iterations."""
    n = 120
    diag, edges, off = R.dense_covisibility_system(n)
    exact = np.linalg.solve(R.dense_matrix(diag, edges, off), _rhs(n))
    jac, it_jac = R.pcg(diag, edges, off, _rhs(n), 1e-10)
    clusters = R.cluster_inverses(diag, edges, off, 16)
    clu, it_clu = R.pcg(diag, edges, off, _rhs(n), 1e-10, clusters=clusters)
    scale = np.linalg.norm(exact)
    assert np.linalg.norm(jac.reshape(-1) - exact) < 1e-6 * scale
    assert np.linalg.norm(clu.reshape(-1) - exact) < 1e-6 * scale
    assert it_clu < it_jac


def test_coarse_corrected_pcg_matches_cholesky():
    """rig_ba_solver.rs:5279-5320 on the 600-node chain: both levels reach the
    dense solution and the two-level solve takes under two thirds of the
    iterations. The reference notes 33 against 58 iterations; this
    restatement takes exactly those."""
    n = 600
    diag, edges, off = R.chain_system(n)
    exact = np.linalg.solve(R.dense_matrix(diag, edges, off), _rhs(n))
    clusters = R.cluster_inverses(diag, edges, off, 32)
    coarse = R.coarse_inverse(diag, edges, off, n, 32)
    one, it_one = R.pcg(diag, edges, off, _rhs(n), 1e-8, clusters=clusters)
    two, it_two = R.pcg(diag, edges, off, _rhs(n), 1e-8, clusters=clusters, coarse=coarse)
    scale = np.linalg.norm(exact)
    assert np.linalg.norm(one.reshape(-1) - exact) < 1e-4 * scale
    assert np.linalg.norm(two.reshape(-1) - exact) < 1e-4 * scale
    assert 3 * it_two < 2 * it_one
    assert (it_one, it_two) == (58, 33)


def test_edge_product_is_the_dense_product():
    """The edge-list product (B x_a into row b, B^T x_b into row a) equals the
    assembled matrix's product."""
    rng = np.random.default_rng(3)
    diag, edges, off = R.chain_system(40, k=3)
    off = off + rng.normal(size=off.shape) * 0.1  # non-symmetric blocks
    x = rng.normal(size=(40, 6))
    y = R.edge_product(diag, edges, off, x)
    assert np.allclose(y.reshape(-1), R.dense_matrix(diag, edges, off) @ x.reshape(-1), rtol=1e-13, atol=1e-12)


def test_zero_rhs_cap_and_breakdown():
    diag, edges, off = R.chain_system(50)
    x, it = R.pcg(diag, edges, off, np.zeros(300), 1e-8)
    assert it == 0 and not x.any()
    x, it = R.pcg(diag, edges, off, _rhs(50), 1e-30, max_iterations=7)
    assert it == 7
    bad = diag.copy()
    bad[3] = -np.eye(6)
    assert R.pcg(bad, edges, off, _rhs(50), 1e-8) is None
