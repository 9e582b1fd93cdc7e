"""CPU restatement (numpy, FP64) of the COLMAP-port rig solver's iterative
reduced-camera solve, the computation behind `ReducedSolveAccelerator::solve`
(pipelines/slam/src/ba_accel.rs:46-157).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
solver (`vba_reduced_solve`). The product never imports it.

Conventions follow the reference's `ReducedCameraSystem`:
- `diag[i]` is node i's damped 6x6 diagonal block;
- edge e couples nodes `edges[e] = (a, b)`, a < b, sorted, and `offdiag[e]`
  is its (row b, column a) block B, so the product adds B x_a to row b and
  B^T x_b to row a;
- `clusters` is a list of `(first node, dense inverse)` over runs of
  consecutive nodes, `coarse` is `(group, leading nodes covered, inverse of
  the aggregated matrix)`.
Blocks here are numpy row-major 6x6 arrays; the C ABI takes them
column-major, as nalgebra stores them.
"""
from __future__ import annotations

import numpy as np


def edge_product(diag, edges, offdiag, x):
    """y = A x for the block system given by its diagonal blocks and its edge
    list (rig_ba_solver.rs:3020-3050: each block read once, contributing
    B x_a to row b and B^T x_b to row a)."""
    y = np.einsum("nij,nj->ni", diag, x)
    if len(edges):
        a = edges[:, 0].astype(np.int64)
        b = edges[:, 1].astype(np.int64)
        np.add.at(y, b, np.einsum("eij,ej->ei", offdiag, x[a]))
        np.add.at(y, a, np.einsum("eji,ej->ei", offdiag, x[b]))
    return y


def dense_matrix(diag, edges, offdiag):
    """The full (6n)^2 matrix of the block system."""
    n = len(diag)
    A = np.zeros((6 * n, 6 * n))
    for i in range(n):
        A[6 * i:6 * i + 6, 6 * i:6 * i + 6] = diag[i]
    for (a, b), blk in zip(edges, offdiag):
        A[6 * b:6 * b + 6, 6 * a:6 * a + 6] = blk
        A[6 * a:6 * a + 6, 6 * b:6 * b + 6] = blk.T
    return A


def cluster_inverses(diag, edges, offdiag, k):
    """Inverses of the matrix restricted to runs of k consecutive nodes
    (cluster_inverses, rig_ba_solver.rs:1112-1166): diagonal blocks plus the
    edges inside a run; None when a run is not positive definite."""
    n = len(diag)
    out = []
    for s in range(0, n, k):
        m = min(k, n - s)
        M = np.zeros((6 * m, 6 * m))
        for i in range(m):
            M[6 * i:6 * i + 6, 6 * i:6 * i + 6] = diag[s + i]
        out.append((s, M))
    for (a, b), blk in zip(edges, offdiag):
        a, b = int(a), int(b)
        if a // k != b // k:
            continue
        M = out[a // k][1]
        ia, ib = a % k, b % k
        M[6 * ib:6 * ib + 6, 6 * ia:6 * ia + 6] = blk
        M[6 * ia:6 * ia + 6, 6 * ib:6 * ib + 6] = blk.T
    res = []
    for s, M in out:
        try:
            L = np.linalg.cholesky(M)
        except np.linalg.LinAlgError:
            return None
        Li = np.linalg.inv(L)
        res.append((s, Li.T @ Li))
    return res


def coarse_inverse(diag, edges, offdiag, n_frames, group, max_aggregates=256):
    """The coarse level (coarse_correction, rig_ba_solver.rs:1235-1275): the
    matrix restricted to one shared 6-vector per aggregate of `group`
    consecutive nodes (at most `max_aggregates` of them), returned as
    (group, n_frames, inverse); None for fewer than two aggregates or a
    matrix that is not positive definite."""
    group = max(group, -(-n_frames // max_aggregates))
    m = -(-n_frames // group)
    if m < 2:
        return None
    A = np.zeros((6 * m, 6 * m))
    for i in range(n_frames):
        g = i // group * 6
        A[g:g + 6, g:g + 6] += diag[i]
    for (ea, eb), blk in zip(edges, offdiag):
        ea, eb = int(ea), int(eb)
        if ea >= n_frames or eb >= n_frames:
            continue
        ga, gb = ea // group * 6, eb // group * 6
        A[gb:gb + 6, ga:ga + 6] += blk
        A[ga:ga + 6, gb:gb + 6] += blk.T
    try:
        L = np.linalg.cholesky(A)
    except np.linalg.LinAlgError:
        return None
    Li = np.linalg.inv(L)
    return group, n_frames, Li.T @ Li


def _block_inverse(d):
    L = np.linalg.cholesky(d)
    Li = np.linalg.inv(L)
    return Li.T @ Li


def pcg(diag, edges, offdiag, rhs, tol, max_iterations=500, clusters=None, coarse=None):
    """pcg_core (rig_ba_solver.rs:1406-1490) on the edge-list product:
    block-Jacobi, or the cluster inverses when given, plus the additive
    coarse correction when given. Returns (x, iterations), or None when a
    diagonal block is not positive definite or p^T A p is not positive and
    finite."""
    diag = np.asarray(diag, dtype=np.float64)
    offdiag = np.asarray(offdiag, dtype=np.float64)
    edges = np.asarray(edges)
    n = len(diag)
    try:
        precond = np.stack([_block_inverse(d) for d in diag])
    except np.linalg.LinAlgError:
        return None

    def apply(r):
        if clusters is None:
            z = np.einsum("nij,nj->ni", precond, r)
        else:
            z = np.zeros_like(r)
            for s, inv in clusters:
                m = inv.shape[0] // 6
                z[s:s + m] = (inv @ r[s:s + m].reshape(-1)).reshape(m, 6)
        if coarse is not None:
            group, n_frames, cinv = coarse
            mm = cinv.shape[0] // 6
            rc = np.zeros((mm, 6))
            for i in range(n_frames):
                rc[i // group] += r[i]
            yc = (cinv @ rc.reshape(-1)).reshape(mm, 6)
            for i in range(n_frames):
                z[i] += yc[i // group]
        return z

    b = np.asarray(rhs, dtype=np.float64).reshape(n, 6)
    b_norm = np.sqrt(np.sum(b * b))
    x = np.zeros((n, 6))
    if b_norm == 0.0:
        return x, 0
    r = b.copy()
    z = apply(r)
    p = z.copy()
    rz = np.sum(r * z)
    iterations = 0
    for _ in range(max_iterations):
        iterations += 1
        ap = edge_product(diag, edges, offdiag, p)
        pap = np.sum(p * ap)
        if not np.isfinite(pap) or pap <= 0.0:
            return None
        alpha = rz / pap
        x += alpha * p
        r -= alpha * ap
        if np.sqrt(np.sum(r * r)) <= tol * b_norm:
            break
        z = apply(r)
        rz_next = np.sum(r * z)
        beta = rz_next / rz
        rz = rz_next
        p = z + beta * p
    return x, iterations


# ---- the reference's test systems (rig_ba_solver.rs iterative_bench) ----

def dense_covisibility_system(n):
    """A city block revisited in four passes: each node couples to its
    +-150 capture-order neighbours and to 100 nodes of the other passes
    (dense_covisibility_system, rig_ba_solver.rs:5161-5200). Returns
    (diag, edges, offdiag) with edges (j, i), j < i, holding the (row i,
    column j) block."""
    cols = [dict() for _ in range(n)]

    def add(a, b, k):
        lo, hi = min(a, b), max(a, b)
        if lo != hi:
            r = np.arange(6)[:, None]
            c = np.arange(6)[None, :]
            cols[lo][hi] = ((r * 7 + c * 3 + lo + hi) % 11).astype(np.float64) * 1e-3 * k

    q = n // 4
    for i in range(n):
        for d in range(1, 151):
            if i + d < n:
                add(i, i + d, 1.0)
        place = i % q
        for k in range(100):
            add(i, ((i // q + 1 + k % 3) % 4) * q + (place + k // 3) % q, 0.5)
    diag = np.stack([np.eye(6) * 10.0 + np.diag(np.full(6, float(i % 5))) for i in range(n)])
    edges, offdiag = [], []
    for j in range(n):
        for i in sorted(cols[j]):
            edges.append((j, i))
            offdiag.append(cols[j][i])
    return diag, np.array(edges, dtype=np.uint32), np.stack(offdiag)


def chain_system(n, k=8):
    """A long open capture: each node couples to its k capture-order
    neighbours on each side as a graph Laplacian plus a little diagonal
    damping (chain_system, rig_ba_solver.rs:5255-5275)."""
    diag = np.stack([np.diag([1e-4 * (1 + (i + r) % 3) for r in range(6)]) for i in range(n)])
    edges, offdiag = [], []
    for a in range(n):
        for b in range(a + 1, min(a + k + 1, n)):
            w = 1.0 + ((a * 7 + b * 3) % 5) * 0.1
            diag[a] += np.eye(6) * w
            diag[b] += np.eye(6) * w
            edges.append((a, b))
            offdiag.append(-w * np.eye(6))
    return diag, np.array(edges, dtype=np.uint32), np.stack(offdiag)
