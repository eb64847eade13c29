"""ctypes binding of the FP64 CPU oracle (oracle/_build/libqgt_oracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libqgt_oracle.so")

GATE = dict(I=0, X=1, Y=2, Z=3, H=4, S=5, T=6, RX=7, RY=8, RZ=9, CNOT=10, CY=11,
            CZ=12, SWAP=13, U1=15, PHASE=19, SDG=26, TDG=27, SX=29)

OP_DTYPE = np.dtype([("type", np.int32), ("q0", np.int32), ("q1", np.int32),
                     ("param", np.int32), ("angle", np.float64)], align=True)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.oracle_uniform_angles.argtypes = [ctypes.c_uint64, ctypes.c_size_t, vp]
        L.oracle_hea_num_ops.argtypes = [ip, ip]
        L.oracle_hea_ops.argtypes = [ip, ip, vp, vp]
        L.oracle_num_params.argtypes = [ip, vp]
        L.oracle_run.argtypes = [ip, ip, vp, vp]
        L.oracle_derivative_states.argtypes = [ip, ip, vp, vp, vp]
        L.oracle_derivative_states_shift.argtypes = [ip, ip, vp, vp]
        L.oracle_derivative_subset.argtypes = [ip, ip, vp, ip, vp, vp, vp]
        L.oracle_qgt_pairwise.argtypes = [ctypes.c_size_t, ip, vp, vp, vp]
        L.oracle_qgt_elementwise.argtypes = [ctypes.c_size_t, ip, vp, vp, vp]
        L.oracle_qgt.argtypes = [ip, ip, vp, vp]
        L.oracle_qgt_hoisted.argtypes = [ip, ip, vp, vp]
        L.oracle_qgt_block.argtypes = [ip, ip, vp, ip, vp, ip, vp, vp]
        L.oracle_derivative_states_timed.argtypes = [ip, ip, vp, ip, vp, vp, vp]
        L.oracle_qgt_loop_time.argtypes = [ctypes.c_size_t, ip, vp, vp, ip, ip, ip, ip, vp, vp]
        L.oracle_dense_unitary.argtypes = [ip, ip, vp, vp]
        L.oracle_pauli_expectation.argtypes = [ip, vp, ctypes.c_char_p]
        L.oracle_pauli_expectation.restype = dp
        L.oracle_tfim_apply.argtypes = [ip, dp, dp, vp, vp]
        L.oracle_tfim_energy.argtypes = [ip, dp, dp, vp]
        L.oracle_tfim_energy.restype = dp
        L.oracle_tfim_gradient.argtypes = [ip, ip, vp, dp, dp, vp, vp]
        L.oracle_cholesky_solve.argtypes = [ip, vp, vp, vp]
        L.oracle_pcg.argtypes = [ip, vp, vp, vp, dp, ip, vp, vp, ip, vp]
        L.oracle_hermitian_eig.argtypes = [ip, vp, vp, vp]
        L.oracle_regularized_natural_gradient.argtypes = [ip, vp, vp, vp, dp, dp, ip, ip, dp]
        L.oracle_regularized_natural_gradient.restype = dp
        L.oracle_mlp_num_params.argtypes = [ip, ip, ip, ip]
        L.oracle_mlp_num_params.restype = ctypes.c_size_t
        L.oracle_mlp_forward.argtypes = [ip, ip, ip, ip, ip, vp, vp, vp]
        L.oracle_kaporin.argtypes = [ip, ip, vp, vp, vp]
        L.oracle_kaporin.restype = dp
        L.oracle_mlp_loss_grad.argtypes = [ip, ip, ip, ip, ip, ip, vp, vp, vp, vp]
        L.oracle_mlp_loss_grad.restype = dp
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def uniform_angles(seed, count):
    out = np.empty(count, np.float64)
    lib().oracle_uniform_angles(seed, count, _p(out))
    return out


def hea_ops(n, layers, theta):
    nops = lib().oracle_hea_num_ops(n, layers)
    ops = np.zeros(nops, OP_DTYPE)
    th = np.ascontiguousarray(theta, np.float64)
    lib().oracle_hea_ops(n, layers, _p(th), _p(ops))
    return ops


def make_ops(gates):
    """gates: list of (name, q0, q1, param, angle)."""
    ops = np.zeros(len(gates), OP_DTYPE)
    for i, (name, q0, q1, param, angle) in enumerate(gates):
        ops[i] = (GATE[name], q0, q1, param, angle)
    return ops


def num_params(ops):
    return int(np.sum(ops["param"] >= 0))


def run(n, ops):
    psi = np.empty(1 << n, np.complex128)
    assert lib().oracle_run(n, len(ops), _p(ops), _p(psi)) == 0
    return psi


def derivative_states(n, ops):
    P = num_params(ops)
    psi = np.empty(1 << n, np.complex128)
    D = np.empty((P, 1 << n), np.complex128)
    assert lib().oracle_derivative_states(n, len(ops), _p(ops), _p(psi), _p(D)) == 0
    return psi, D


def derivative_states_shift(n, ops):
    P = num_params(ops)
    D = np.empty((P, 1 << n), np.complex128)
    assert lib().oracle_derivative_states_shift(n, len(ops), _p(ops), _p(D)) == 0
    return D


def qgt(n, ops):
    P = num_params(ops)
    Q = np.empty((P, P), np.complex128)
    assert lib().oracle_qgt(n, len(ops), _p(ops), _p(Q)) == 0
    return Q


def qgt_hoisted(n, ops):
    P = num_params(ops)
    Q = np.empty((P, P), np.complex128)
    assert lib().oracle_qgt_hoisted(n, len(ops), _p(ops), _p(Q)) == 0
    return Q


def qgt_from_states(psi, D, variant="elementwise"):
    P, dim = D.shape
    Q = np.empty((P, P), np.complex128)
    fn = lib().oracle_qgt_pairwise if variant == "pairwise" else lib().oracle_qgt_elementwise
    fn(dim, P, _p(np.ascontiguousarray(psi)), _p(np.ascontiguousarray(D)), _p(Q))
    return Q


def derivative_states_timed(n, ops, threads, psi=None, D=None):
    """Derivative states on `threads` threads (<= 0: default), into psi / D when given.
    Returns (psi, D, seconds)."""
    P = num_params(ops)
    if psi is None:
        psi = np.empty(1 << n, np.complex128)
    if D is None:
        D = np.empty((P, 1 << n), np.complex128)
    sec = ctypes.c_double(0.0)
    assert lib().oracle_derivative_states_timed(n, len(ops), _p(ops), threads, _p(psi), _p(D), ctypes.byref(sec)) == 0
    return psi, D, sec.value


def qgt_loop_time(psi, D, variant, threads, rows, reps):
    """Rows [0, rows) of the per-element loop ("elementwise": mu <= nu, three dot
    products per element) or of the P x P loop ("pairwise") over the given states,
    `reps` times.  Returns ([seconds per repetition], elements per repetition)."""
    P, dim = D.shape
    el = ctypes.c_long(0)
    t = np.zeros(reps, np.float64)
    v = 1 if variant == "pairwise" else 0
    assert lib().oracle_qgt_loop_time(dim, P, _p(psi), _p(D), v, threads, rows, reps, _p(t), ctypes.byref(el)) == 0
    return list(t), el.value


def qgt_block(n, ops, rows, cols):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    Q = np.empty((len(rows), len(cols)), np.complex128)
    assert lib().oracle_qgt_block(n, len(ops), _p(ops), len(rows), _p(rows), len(cols), _p(cols), _p(Q)) == 0
    return Q


def dense_unitary(n, ops):
    U = np.empty((1 << n, 1 << n), np.complex128)
    assert lib().oracle_dense_unitary(n, len(ops), _p(ops), _p(U)) == 0
    return U


def pauli_expectation(n, psi, pauli):
    return lib().oracle_pauli_expectation(n, _p(np.ascontiguousarray(psi, np.complex128)), pauli.encode())


def tfim_apply(n, J, h, psi):
    out = np.empty_like(psi)
    lib().oracle_tfim_apply(n, J, h, _p(psi), _p(out))
    return out


def tfim_energy(n, J, h, psi):
    return lib().oracle_tfim_energy(n, J, h, _p(np.ascontiguousarray(psi, np.complex128)))


def tfim_gradient(n, ops, J, h):
    P = num_params(ops)
    e = np.zeros(1, np.float64)
    g = np.empty(P, np.float64)
    assert lib().oracle_tfim_gradient(n, len(ops), _p(ops), J, h, _p(e), _p(g)) == 0
    return float(e[0]), g


def cholesky_solve(A, b):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    rc = lib().oracle_cholesky_solve(len(b), _p(A), _p(b), _p(x))
    return x if rc == 0 else None


def pcg(A, b, tol=1e-8, maxit=10000, d=None, U=None):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    res = np.zeros(1, np.float64)
    dd = None if d is None else np.ascontiguousarray(d, np.float64)
    UU = None if U is None else np.ascontiguousarray(U, np.float64)
    r = 0 if U is None else UU.shape[1]
    it = lib().oracle_pcg(len(b), _p(A), _p(b), _p(x), tol, maxit,
                          None if dd is None else _p(dd), None if UU is None else _p(UU), r, _p(res))
    return x, it, float(res[0])


def hermitian_eig(A):
    A = np.ascontiguousarray(A, np.complex128)
    P = A.shape[0]
    ev = np.empty(P, np.float64)
    V = np.empty((P, P), np.complex128)
    lib().oracle_hermitian_eig(P, _p(A), _p(ev), _p(V))
    return ev, V


def regularized_natural_gradient(G, g, lam=1e-4, cond=1e8, adaptive=True, pinv=True, cutoff=1e-10):
    G = np.ascontiguousarray(G, np.complex128)
    g = np.ascontiguousarray(g, np.complex128)
    x = np.empty_like(g)
    used = lib().oracle_regularized_natural_gradient(len(g), _p(G), _p(g), _p(x), lam, cond,
                                                     int(adaptive), int(pinv), cutoff)
    return x, used


# ------------------------------------------------------------------ learned preconditioner
def mlp_num_params(F, H, nh, O):
    return int(lib().oracle_mlp_num_params(F, H, nh, O))


def mlp_forward(F, H, nh, O, params, X):
    X = np.ascontiguousarray(X, np.float64)
    params = np.ascontiguousarray(params, np.float64)
    Y = np.zeros((X.shape[0], O))
    lib().oracle_mlp_forward(F, H, nh, O, X.shape[0], _p(params), _p(X), _p(Y))
    return Y


def kaporin(P, r, y, As, grad=True):
    y = np.ascontiguousarray(y, np.float64)
    As = np.ascontiguousarray(As, np.float64)
    dy = np.zeros_like(y) if grad else None
    loss = lib().oracle_kaporin(P, r, _p(y), _p(As), _p(dy) if grad else None)
    return (loss, dy) if grad else loss


def mlp_loss_grad(F, H, nh, P, r, params, X, As, grad=True):
    X = np.ascontiguousarray(X, np.float64)
    As = np.ascontiguousarray(As, np.float64)
    params = np.ascontiguousarray(params, np.float64)
    g = np.zeros_like(params) if grad else None
    loss = lib().oracle_mlp_loss_grad(F, H, nh, P, r, X.shape[0], _p(params), _p(X), _p(As), _p(g) if grad else None)
    return (loss, g) if grad else loss
