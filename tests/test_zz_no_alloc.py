"""The no-allocation solve (qgtl_solve_f64_noalloc) through the C ABI against
the allocating one (qgtl_solve_f64) on the same inputs: two calls in a row,
the second smaller, for every method; plus the size limit and proof that the
calls made no device or pinned-host allocation.

Bar: the two entry points run the same kernels on the same inputs, so x agrees
to 1e-13 relative (with equal iteration counts); as a sanity check both agree
with numpy's FP64 solve (1e-10 for Cholesky; 1e-5 for the iterative methods
run to a relative residual of 1e-9 on systems of condition number ~1e4).
"""
import ctypes

import numpy as np
import pytest

import quantum_geometric_tensor_amd as Q

pytestmark = pytest.mark.gpu

LAM = 1e-3


def _systems(seed, batch, P, rank):
    """Metric-like systems: A = Re(V V^H) / P, PSD with a decaying spectrum."""
    rng = np.random.default_rng(seed)
    A = np.empty((batch, P, P))
    for s in range(batch):
        V = rng.standard_normal((P, P)) * np.exp(-np.arange(P) / (P / 6.0))
        A[s] = V @ V.T / P
    b = rng.standard_normal((batch, P))
    d = 1.0 / (np.einsum("sii->si", A) + LAM)
    U = rng.standard_normal((batch, P, rank)) * 0.01
    return A, b, d, U


def _noalloc(A, b, method, tol, d=None, U=None):
    L = Q.lib()
    fn = L.qgtl_solve_f64_noalloc
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                   ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p]
    batch, P = b.shape
    A = np.ascontiguousarray(A)
    b = np.ascontiguousarray(b)
    x = np.zeros((batch, P))
    info = (Q.SolveInfo * batch)()
    dd = None if d is None else np.ascontiguousarray(d)
    UU = None if U is None else np.ascontiguousarray(U)
    rank = 0 if UU is None else UU.shape[-1]
    rc = fn(P, batch, A.ctypes.data, b.ctypes.data, LAM, method, tol, 10000,
            None if dd is None else dd.ctypes.data, None if UU is None else UU.ctypes.data, rank, x.ctypes.data,
            ctypes.cast(info, ctypes.c_void_p))
    return rc, x, [(i.status, i.iterations, i.rel_residual) for i in info]


def _alloc_calls():
    fn = Q.lib().qgtl_debug_alloc_calls
    fn.restype = ctypes.c_ulonglong
    return fn()


CASES = [(Q.SOLVE_CHOLESKY, 0.0), (Q.SOLVE_CG, 1e-9), (Q.SOLVE_JACOBI_PCG, 1e-9), (Q.SOLVE_PCG, 1e-9)]


@pytest.mark.parametrize("method,tol", CASES)
def test_noalloc_solve_matches_solve_twice(method, tol):
    Q.lib()
    Q.solve(np.eye(4), np.ones(4))          # library initialised (streams, scratch address)
    for call, (batch, P, rank) in enumerate([(6, 192, 8), (2, 40, 3)]):
        A, b, d, U = _systems(call, batch, P, rank)
        use_d = method == Q.SOLVE_PCG
        x_ref, info_ref = Q.solve(A, b, LAM, method, tol=tol, d=d if use_d else None, U=U if use_d else None)
        before = _alloc_calls()
        rc, x, info = _noalloc(A, b, method, tol, d if use_d else None, U if use_d else None)
        after = _alloc_calls()
        print(f"noalloc call {call}: method {method}, batch {batch}, P {P}: rc {rc}, "
              f"iterations {[i[1] for i in info]}", flush=True)
        assert rc == 0
        assert after == before, "the no-allocation solve allocated device or pinned memory"
        rel = np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref))
        assert rel <= 1e-13, rel
        for (st, it, res), ref in zip(info, info_ref):
            assert st == 0 and it == ref["iterations"]
        x_np = np.stack([np.linalg.solve(A[s] + LAM * np.eye(P), b[s]) for s in range(batch)])
        err = np.max(np.abs(x - x_np)) / np.max(np.abs(x_np))
        assert err <= (1e-10 if method == Q.SOLVE_CHOLESKY else 1e-5), err


def test_noalloc_solve_rejects_oversize():
    Q.lib()
    P, batch = 1024, 5                       # 5 x 8 MiB of matrices: beyond the 32 MiB scratch
    A = np.zeros((batch, P, P))
    b = np.zeros((batch, P))
    before = _alloc_calls()
    rc, _, _ = _noalloc(A, b, Q.SOLVE_CHOLESKY, 0.0)
    assert _alloc_calls() == before          # no allocation on the error path either
    assert rc == -1                          # QGT_ERROR_INVALID_ARGUMENT
    msg = Q.lib().qgtl_last_error().decode()
    assert "scratch" in msg and "maximum" in msg, msg
