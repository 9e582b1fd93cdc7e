"""GPU parity of vba_reduced_solve (ReducedSolveAccelerator::solve,
pipelines/slam/src/ba_accel.rs:46-157) against the PCG oracle
(oracle/reduced_pcg.py, a restatement of rig_ba_solver.rs:1406-1490).

Both run FP64 arithmetic on the same FP32 off-diagonal blocks, so they take
the same iterations and their steps agree to rounding: within 1e-9 of the
step's norm. Against the dense solve they agree to the PCG tolerance."""
import numpy as np
import pytest

import reduced_pcg as R

pytestmark = pytest.mark.gpu

STEP_RTOL = 1e-9


def _rhs(n):
    return (np.arange(6 * n) * 13 % 17).astype(np.float64) - 8.0


def _f32(off):
    return np.asarray(off, dtype=np.float32).astype(np.float64)


def _compare(device, diag, edges, off, rhs, tol, clusters=None, coarse=None, max_iterations=500):
    off = _f32(off)
    ref = R.pcg(diag, edges, off, rhs, tol, max_iterations, clusters=clusters, coarse=coarse)
    out = device.reduced_solve(diag, edges, off, rhs, tol, max_iterations, clusters=clusters, coarse=coarse)
    assert out is not None, device.last_error()
    x, it = out
    xr, itr = ref
    assert it == itr, (it, itr)
    assert np.linalg.norm(x - xr) <= STEP_RTOL * np.linalg.norm(xr), np.linalg.norm(x - xr) / np.linalg.norm(xr)
    return x, it


def test_covisibility_jacobi_and_clusters(device):
    """rig_ba_solver.rs:5202-5250's system: block-Jacobi and runs of 16."""
    n = 120
    diag, edges, off = R.dense_covisibility_system(n)
    exact = np.linalg.solve(R.dense_matrix(diag, edges, _f32(off)), _rhs(n))
    x1, it1 = _compare(device, diag, edges, off, _rhs(n), 1e-10)
    clusters = R.cluster_inverses(diag, edges, _f32(off), 16)
    x2, it2 = _compare(device, diag, edges, off, _rhs(n), 1e-10, clusters=clusters)
    assert it2 < it1
    for x in (x1, x2):
        assert np.linalg.norm(x.reshape(-1) - exact) < 1e-6 * np.linalg.norm(exact)


def test_chain_two_level(device):
    """rig_ba_solver.rs:5279-5320's chain with runs of 32 and the coarse level."""
    n = 600
    diag, edges, off = R.chain_system(n)
    clusters = R.cluster_inverses(diag, edges, _f32(off), 32)
    coarse = R.coarse_inverse(diag, edges, _f32(off), n, 32)
    _, it_one = _compare(device, diag, edges, off, _rhs(n), 1e-8, clusters=clusters)
    _, it_two = _compare(device, diag, edges, off, _rhs(n), 1e-8, clusters=clusters, coarse=coarse)
    assert 3 * it_two < 2 * it_one


def _random_system(n, per_node, seed, extras=0):
    """A diagonally dominant SPD block system with `per_node` forward edges a
    node and non-symmetric blocks."""
    rng = np.random.default_rng(seed)
    edges, off = [], []
    for a in range(n):
        bs = np.unique(rng.integers(a + 1, min(n, a + 200) + 1, size=per_node)) if a + 1 < n else []
        for b in bs:
            if b < n:
                edges.append((a, int(b)))
                off.append(rng.normal(size=(6, 6)) * 0.05)
    edges = np.array(edges, dtype=np.uint32)
    off = np.stack(off)
    diag = np.zeros((n, 6, 6))
    for (a, b), B in zip(edges, off):
        s = np.abs(B).sum()
        diag[a] += np.eye(6) * s
        diag[b] += np.eye(6) * s
    for i in range(n):
        G = rng.normal(size=(6, 6)) * 0.1
        diag[i] += G @ G.T + np.eye(6) * 0.5
    return diag, edges, off


@pytest.mark.parametrize("precond", ["jacobi", "clusters", "two-level"])
def test_random_system(device, precond):
    n = 3000
    diag, edges, off = _random_system(n, 12, seed=5)
    rhs = np.random.default_rng(6).normal(size=6 * n)
    clusters = coarse = None
    if precond != "jacobi":
        clusters = R.cluster_inverses(diag, edges, _f32(off), 32)
    if precond == "two-level":
        coarse = R.coarse_inverse(diag, edges, _f32(off), n - 7, 32)  # leading nodes only
    _compare(device, diag, edges, off, rhs, 1e-6, clusters=clusters, coarse=coarse)


def test_edge_cases(device):
    diag, edges, off = R.chain_system(50)
    x, it = device.reduced_solve(diag, edges, off, np.zeros(300), 1e-8)
    assert it == 0 and not x.any()
    _compare(device, diag, edges, off, _rhs(50), 1e-30, max_iterations=7)
    bad = diag.copy()
    bad[3] = -np.eye(6)
    assert device.reduced_solve(bad, edges, off, _rhs(50), 1e-8) is None
    assert "positive" in device.last_error()
    unsorted = edges[::-1].copy()
    assert device.reduced_solve(diag, unsorted, off[::-1], _rhs(50), 1e-8) is None
    assert "sorted" in device.last_error()


def test_repeatable_and_structure_reuse(device):
    n = 800
    diag, edges, off = _random_system(n, 10, seed=9)
    rhs = _rhs(n)
    a = device.reduced_solve(diag, edges, off, rhs, 1e-8, structure_id=77)
    b = device.reduced_solve(diag, edges, off, rhs, 1e-8, structure_id=77)  # cached edge index
    c = device.reduced_solve(diag, edges, off, rhs, 1e-8, structure_id=0)
    assert a[1] == b[1] == c[1]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[0], c[0])
