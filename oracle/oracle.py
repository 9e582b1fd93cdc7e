"""ctypes binding of the CPU FP64 oracle (oracle/ba_oracle.h).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
product (libvba.so and its Python package) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libba_oracle.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def library():
    global _lib
    if _lib is not None:
        return _lib
    import vba.abi as abi  # struct definitions shared with the C header

    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P64 = abi.P64
    lib.oracle_optimize.restype = C.c_int32
    lib.oracle_optimize.argtypes = [C.POINTER(abi.Problem), C.POINTER(abi.Options), P64, P64,
                                    C.POINTER(abi.IterationStats), C.c_uint32, C.POINTER(abi.Result),
                                    C.c_int32]
    lib.oracle_optimize_weighted.restype = C.c_int32
    lib.oracle_optimize_weighted.argtypes = [C.POINTER(abi.Problem), C.POINTER(abi.Options), P64, C.c_uint64, P64,
                                             P64, C.POINTER(abi.IterationStats), C.c_uint32, C.POINTER(abi.Result),
                                             C.c_int32]
    lib.oracle_optimize_intrinsics.restype = C.c_int32
    lib.oracle_optimize_intrinsics.argtypes = [C.POINTER(abi.Problem), C.POINTER(abi.Options), P64, P64,
                                               C.POINTER(abi.Camera), C.POINTER(abi.IterationStats), C.c_uint32,
                                               C.POINTER(abi.Result), C.c_int32]
    lib.oracle_default_gnc_options.restype = None
    lib.oracle_default_gnc_options.argtypes = [C.POINTER(abi.GncOptions)]
    lib.oracle_optimize_gnc.restype = C.c_int32
    lib.oracle_optimize_gnc.argtypes = [C.POINTER(abi.Problem), C.POINTER(abi.Options), C.POINTER(abi.GncOptions),
                                        P64, P64, P64, C.POINTER(abi.GncResult), C.c_int32]
    lib.oracle_gnc_state_weights.restype = C.c_int32
    lib.oracle_gnc_state_weights.argtypes = [C.POINTER(abi.GncOptions), C.c_double, C.c_uint32, P64, C.c_uint32,
                                             P64, P64]
    lib.oracle_estimate_scale_mad.restype = C.c_int32
    lib.oracle_estimate_scale_mad.argtypes = [P64, C.c_uint32, C.c_double, P64]
    lib.oracle_camera_project.restype = C.c_int32
    lib.oracle_camera_project.argtypes = [C.POINTER(abi.Camera), P64, P64]
    lib.oracle_mono_project.restype = C.c_int32
    lib.oracle_mono_project.argtypes = [C.POINTER(abi.Camera), P64, P64, P64]
    lib.oracle_robust.restype = None
    lib.oracle_robust.argtypes = [C.c_uint32, C.c_double, C.c_double, P64, P64]
    lib.oracle_linearize.restype = C.c_int32
    lib.oracle_linearize.argtypes = [C.POINTER(abi.Problem), C.c_uint32, C.c_double, P64, P64, P64, P64,
                                     C.POINTER(C.c_uint8)]
    lib.oracle_cost.restype = C.c_int32
    lib.oracle_cost.argtypes = [C.POINTER(abi.Problem), C.c_uint32, C.c_double, P64, C.POINTER(C.c_uint64)]
    for name in ("oracle_se3_exp", "oracle_se3_log", "oracle_camera_center"):
        getattr(lib, name).restype = None
        getattr(lib, name).argtypes = [P64, P64]
    lib.oracle_se3_compose.restype = None
    lib.oracle_se3_compose.argtypes = [P64, P64, P64]
    ns = C.POINTER(NormalSystem)
    lib.oracle_ns_reduce_dense.restype = None
    lib.oracle_ns_reduce_dense.argtypes = [ns, C.c_double, P64, P64]
    lib.oracle_ns_solve_dense.restype = C.c_int32
    lib.oracle_ns_solve_dense.argtypes = [ns, C.c_double, P64, P64]
    lib.oracle_ns_operator.restype = C.c_int32
    lib.oracle_ns_operator.argtypes = [ns, C.c_double, P64, P64, P64, P64, P64]
    lib.oracle_ns_solve_pcg.restype = C.c_int32
    lib.oracle_ns_solve_pcg.argtypes = [ns, C.c_double, C.c_uint32, C.c_double, C.c_double, C.c_int32, P64,
                                        P64, C.POINTER(C.c_uint32), P64, P64]
    lib.oracle_reduced_eliminate.restype = C.c_int32
    lib.oracle_reduced_eliminate.argtypes = [C.POINTER(abi.PointElimination), P64, P64, P64]
    lib.oracle_cholesky.restype = C.c_int32
    lib.oracle_cholesky.argtypes = [C.c_uint32, P64, P64, C.c_int32]
    lib.oracle_solve_normal_equations.restype = C.c_int32
    lib.oracle_solve_normal_equations.argtypes = [C.c_uint32, P64, P64, P64]
    _lib = lib
    return lib


class NormalSystem(C.Structure):
    _fields_ = [
        ("num_poses", C.c_uint32),
        ("diag", C.POINTER(C.c_double)),
        ("b_p", C.POINTER(C.c_double)),
        ("num_landmarks", C.c_uint32),
        ("h_ll", C.POINTER(C.c_double)),
        ("b_l", C.POINTER(C.c_double)),
        ("cross_ptr", C.POINTER(C.c_uint32)),
        ("cross_pose", C.POINTER(C.c_uint32)),
        ("cross", C.POINTER(C.c_double)),
    ]


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def optimize(problem, options, threads: int = 1, cap: int = 1024):
    """oracle_optimize on a vba.BundleProblem; returns vba.SolveOutput."""
    import vba
    import vba.abi as abi

    lib = library()

    def call(view, opts, poses, lms, stats, cap_, res):
        return lib.oracle_optimize(C.byref(view), C.byref(opts), _p(poses), _p(lms), stats, cap_, C.byref(res),
                                   threads)

    return vba.run_solver(call, problem, options, cap)


def optimize_weighted(problem, options, weights, threads: int = 1, cap: int = 1024):
    """oracle_optimize_weighted; `weights` one per observation (mono, then rig)."""
    import vba

    lib = library()
    w = np.ascontiguousarray(weights, np.float64)

    def call(view, opts, poses, lms, stats, cap_, res):
        return lib.oracle_optimize_weighted(C.byref(view), C.byref(opts), _p(w), len(w), _p(poses), _p(lms), stats,
                                            cap_, C.byref(res), threads)

    return vba.run_solver(call, problem, options, cap)


def optimize_intrinsics(problem, options, threads: int = 1, cap: int = 1024):
    """oracle_optimize_intrinsics: returns (vba.SolveOutput, refined abi.Camera)."""
    import vba

    lib = library()
    cam = abi_camera()

    def call(view, opts, poses, lms, stats, cap_, res):
        return lib.oracle_optimize_intrinsics(C.byref(view), C.byref(opts), _p(poses), _p(lms), C.byref(cam), stats,
                                              cap_, C.byref(res), threads)

    return vba.run_solver(call, problem, options, cap), cam


def abi_camera():
    import vba.abi as abi
    return abi.Camera()


def default_gnc_options(**overrides):
    import vba.abi as abi

    g = abi.GncOptions()
    library().oracle_default_gnc_options(C.byref(g))
    for k, v in overrides.items():
        setattr(g, k, v)
    return g


def optimize_gnc(problem, options, gnc, threads: int = 1):
    """oracle_optimize_gnc: (status, GncResult, poses, landmarks, weights)."""
    import vba.abi as abi

    lib = library()
    view = problem.view()
    poses = np.empty_like(problem.poses)
    lms = np.empty_like(problem.landmarks)
    w = np.empty(problem.num_observations)
    res = abi.GncResult()
    st = lib.oracle_optimize_gnc(C.byref(view), C.byref(options), C.byref(gnc), _p(poses), _p(lms), _p(w),
                                 C.byref(res), threads)
    return st, res, poses, lms, w


def gnc_state_weights(gnc, max_residual_sq, anneal_steps, s):
    """GncState::new + anneal() x anneal_steps, then weight(s): (weights, mu, terminal)."""
    s = np.ascontiguousarray(s, np.float64)
    w = np.empty_like(s)
    mu = C.c_double()
    t = library().oracle_gnc_state_weights(C.byref(gnc), max_residual_sq, anneal_steps, _p(s), len(s), _p(w),
                                           C.byref(mu))
    return w, mu.value, bool(t)


def estimate_scale_mad(squared, k):
    s = np.ascontiguousarray(squared, np.float64)
    out = C.c_double()
    ok = library().oracle_estimate_scale_mad(_p(s), len(s), k, C.byref(out))
    return out.value if ok else None


def camera_project(cam, xc):
    lib = library()
    xc = np.ascontiguousarray(xc, np.float64)
    uv = np.zeros(2)
    ok = lib.oracle_camera_project(C.byref(cam), _p(xc), _p(uv))
    return uv if ok else None


def mono_project(cam, xc):
    lib = library()
    xc = np.ascontiguousarray(xc, np.float64)
    uv = np.zeros(2)
    j = np.zeros(6)
    ok = lib.oracle_mono_project(C.byref(cam), _p(xc), _p(uv), _p(j))
    return (uv, j.reshape(2, 3)) if ok else None


def robust(kernel, scale, s):
    lib = library()
    c = np.zeros(1)
    w = np.zeros(1)
    lib.oracle_robust(kernel, scale, s, _p(c), _p(w))
    return float(c[0]), float(w[0])


def linearize(problem, kernel=0, scale=0.0):
    lib = library()
    n = problem.num_observations
    r = np.zeros((n, 2))
    jp = np.zeros((n, 2, 6))
    jl = np.zeros((n, 2, 3))
    w = np.zeros(n)
    valid = np.zeros(n, np.uint8)
    view = problem.view()
    st = lib.oracle_linearize(C.byref(view), kernel, scale, _p(r), _p(jp), _p(jl), _p(w),
                              valid.ctypes.data_as(C.POINTER(C.c_uint8)))
    if st != 0:
        raise ValueError(f"oracle_linearize status {st}")
    return r, jp, jl, w, valid.astype(bool)


def cost(problem, kernel=0, scale=0.0):
    lib = library()
    c = np.zeros(1)
    n = C.c_uint64()
    view = problem.view()
    st = lib.oracle_cost(C.byref(view), kernel, scale, _p(c), C.byref(n))
    if st != 0:
        raise ValueError(f"oracle_cost status {st}")
    return float(c[0]), int(n.value)


def se3_exp(xi):
    out = np.zeros(7)
    library().oracle_se3_exp(_p(np.ascontiguousarray(xi, np.float64)), _p(out))
    return out


def se3_log(pose):
    out = np.zeros(6)
    library().oracle_se3_log(_p(np.ascontiguousarray(pose, np.float64)), _p(out))
    return out


def se3_compose(a, b):
    out = np.zeros(7)
    library().oracle_se3_compose(_p(np.ascontiguousarray(a, np.float64)), _p(np.ascontiguousarray(b, np.float64)),
                                 _p(out))
    return out


def camera_center(pose):
    out = np.zeros(3)
    library().oracle_camera_center(_p(np.ascontiguousarray(pose, np.float64)), _p(out))
    return out


class System:
    """Owner of the arrays behind an oracle_normal_system."""

    def __init__(self, diag, b_p, h_ll, b_l, cross):
        # cross: list per landmark of [(pose, 6x3 array)]
        self.diag = np.ascontiguousarray(np.asarray(diag, np.float64).reshape(-1, 36))
        self.b_p = np.ascontiguousarray(np.asarray(b_p, np.float64).reshape(-1))
        self.h_ll = np.ascontiguousarray(np.asarray(h_ll, np.float64).reshape(-1, 9))
        self.b_l = np.ascontiguousarray(np.asarray(b_l, np.float64).reshape(-1, 3))
        ptr = [0]
        poses, blocks = [], []
        for entries in cross:
            for p, m in entries:
                poses.append(p)
                blocks.append(np.asarray(m, np.float64).reshape(18))
            ptr.append(len(poses))
        self.cross_ptr = np.array(ptr, np.uint32)
        self.cross_pose = np.array(poses if poses else [0], np.uint32)
        self.cross = np.ascontiguousarray(np.array(blocks if blocks else [np.zeros(18)], np.float64))
        self.P = self.diag.shape[0]
        self.L = self.h_ll.shape[0]

    def view(self):
        v = NormalSystem()
        v.num_poses = self.P
        v.diag = _p(self.diag)
        v.b_p = _p(self.b_p)
        v.num_landmarks = self.L
        v.h_ll = _p(self.h_ll)
        v.b_l = _p(self.b_l)
        v.cross_ptr = self.cross_ptr.ctypes.data_as(C.POINTER(C.c_uint32))
        v.cross_pose = self.cross_pose.ctypes.data_as(C.POINTER(C.c_uint32))
        v.cross = _p(self.cross)
        return v

    def reduce_dense(self, lam):
        n = 6 * self.P
        s = np.zeros((n, n))
        rhs = np.zeros(n)
        v = self.view()
        library().oracle_ns_reduce_dense(C.byref(v), lam, _p(s), _p(rhs))
        return s, rhs

    def solve_dense(self, lam):
        dp = np.zeros(6 * self.P)
        dl = np.zeros(3 * self.L)
        v = self.view()
        st = library().oracle_ns_solve_dense(C.byref(v), lam, _p(dp), _p(dl))
        return st, dp, dl

    def operator(self, lam, x=None, r=None):
        n = 6 * self.P
        y = np.zeros(n)
        z = np.zeros(n)
        rhs = np.zeros(n)
        v = self.view()
        xa = np.ascontiguousarray(x, np.float64) if x is not None else None
        ra = np.ascontiguousarray(r, np.float64) if r is not None else None
        st = library().oracle_ns_operator(C.byref(v), lam, _p(xa) if xa is not None else None,
                                          _p(y) if xa is not None else None,
                                          _p(ra) if ra is not None else None,
                                          _p(z) if ra is not None else None, _p(rhs))
        return st, y, z, rhs

    def solve_pcg(self, lam, max_iterations=128, rtol=1e-12, atol=1e-12, accept_at_cap=False):
        dp = np.zeros(6 * self.P)
        dl = np.zeros(3 * self.L)
        it = C.c_uint32()
        rn = np.zeros(1)
        tg = np.zeros(1)
        v = self.view()
        st = library().oracle_ns_solve_pcg(C.byref(v), lam, max_iterations, rtol, atol, int(accept_at_cap),
                                           _p(dp), _p(dl), C.byref(it), _p(rn), _p(tg))
        return st, dp, dl, int(it.value), float(rn[0]), float(tg[0])


def solve_normal_equations(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(b))
    ok = library().oracle_solve_normal_equations(len(b), _p(a), _p(b), _p(x))
    return x if ok else None


def reduced_eliminate(elim):
    """oracle_reduced_eliminate on a vba.PointEliminationInput: (blocks
    (E, 6, 6), run_blocks (R, 6, 6), coarse_sums (m, m, 6, 6) or None), all
    row-major 6x6 here; the CPU rig solver's eliminate_by_output
    (rig_ba_solver.rs:2551-2660) plus the direct couplings, run blocks and
    aggregate sums of ReducedSolveAccelerator::eliminate."""
    lib = library()
    view, keep = elim.view()
    E = len(keep["edges"])
    blocks = np.zeros((max(E, 1), 36))
    runs = np.zeros((max(1, len(keep["run_edges"])), 36))
    m = elim.coarse_aggregates()
    coarse = np.zeros((max(1, m * m), 36))
    st = lib.oracle_reduced_eliminate(C.byref(view), _p(blocks), _p(runs), _p(coarse))
    assert st == 0, st
    rm = lambda a: np.swapaxes(a.reshape(-1, 6, 6), -1, -2).copy()
    return (rm(blocks[:E]), rm(runs[:len(keep["run_edges"])]),
            np.swapaxes(coarse[:m * m].reshape(m, m, 6, 6), -1, -2).copy() if m else None)


def cholesky(a, threads: int = 1):
    """oracle_cholesky: the factor L (lower triangle) of a row-major SPD
    matrix, or None when the factor fails."""
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    out = np.zeros_like(a)
    ok = library().oracle_cholesky(n, _p(a), _p(out), threads)
    return np.tril(out) if ok else None
