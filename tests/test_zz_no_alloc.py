"""vba_dense_solve_noalloc, the dense solve whose device buffers live in the
library's fixed device array, against vba_dense_solve (the same kernels on
buffers it allocates) and the oracle's solve_normal_equations
(pose_graph.rs:2259-2270): bitwise equal to the allocating entry point on two
calls in a row (the second smaller, so stale contents of the larger first
call must not leak into it), on the Cholesky, LU-fallback and singular paths,
and a clear error beyond the maximum size."""
import ctypes as C

import numpy as np
import pytest

import oracle
import vba
from vba import abi

pytestmark = pytest.mark.gpu

SOLVE_RTOL = 1e-10  # blocked MFMA Cholesky / unblocked LU vs nalgebra's order
MAX_N = 1024  # VBA_DENSE_NOALLOC_MAX_N


def spd(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return a @ a.T + n * np.eye(n), rng.standard_normal(n)


def indefinite(n, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = rng.uniform(0.5, 2.0, n) * np.where(np.arange(n) % 3 == 1, -1.0, 1.0)
    return (q * ev) @ q.T, rng.standard_normal(n)


def raw_noalloc(dev, h, b):
    """The C entry point called directly, column-major H as the header says."""
    n = len(b)
    hf = np.asfortranarray(h, np.float64).ravel(order="F")
    bb = np.ascontiguousarray(b, np.float64)
    x = np.full(n, np.nan)
    st = dev.lib.vba_dense_solve_noalloc(dev.ctx, C.c_uint32(n), hf.ctypes.data_as(abi.P64),
                                         bb.ctypes.data_as(abi.P64), x.ctypes.data_as(abi.P64))
    return st, x


def check_pair(dev, h, b):
    st_a, x_a = dev.dense_solve(h, b)
    st_n, x_n = raw_noalloc(dev, h, b)
    assert st_a == st_n
    if st_a == abi.OK:
        assert np.array_equal(x_a, x_n)
        ref = oracle.solve_normal_equations(h, b)
        assert ref is not None
        assert np.abs(x_n - ref).max() <= SOLVE_RTOL * np.abs(ref).max()
    return st_n


@pytest.mark.parametrize("sizes", [(300, 37), (MAX_N, 96)])
def test_noalloc_cholesky_two_calls(device, sizes):
    for k, n in enumerate(sizes):
        h, b = spd(n, 11 + k)
        assert check_pair(device, h, b) == abi.OK


def test_noalloc_lu_fallback_two_calls(device):
    for k, n in enumerate((130, 48)):
        h, b = indefinite(n, 200 + k)
        assert np.linalg.eigvalsh(h).min() < 0
        assert check_pair(device, h, b) == abi.OK


def test_noalloc_after_singular_call(device):
    h, b = spd(60, 5)
    h[9, :] = 0.0
    h[:, 9] = 0.0
    st, _ = raw_noalloc(device, h, b)
    assert st == abi.ERR_SINGULAR_SYSTEM
    assert "singular" in device.last_error()
    # the status word left by the singular call must not leak into the next
    h2, b2 = spd(30, 6)
    assert check_pair(device, h2, b2) == abi.OK


def test_noalloc_rejects_beyond_maximum(device):
    n = MAX_N + 1
    h = np.eye(n)
    b = np.ones(n)
    st, _ = raw_noalloc(device, h, b)
    assert st == abi.ERR_INVALID_ARGUMENT
    assert "exceeds the maximum" in device.last_error()
    # and the entry point still works afterwards
    h2, b2 = spd(20, 7)
    assert check_pair(device, h2, b2) == abi.OK


def test_noalloc_empty_system(device):
    st, x = raw_noalloc(device, np.zeros((0, 0)), np.zeros(0))
    assert st == abi.OK and x.size == 0


def test_noalloc_second_context(device):
    """A second context on the same GPU uses the same array, looked up once."""
    with vba.Device(0) as other:
        h, b = spd(64, 8)
        st1, x1 = raw_noalloc(other, h, b)
        st2, x2 = raw_noalloc(device, h, b)
        assert st1 == st2 == abi.OK
        assert np.array_equal(x1, x2)
