"""quantum_geometric_tensor_amd -- Python view of libqgtl_mi355x.so.

The library is C host code plus hand-written HIP kernels for gfx950 (see
csrc/).  This module only binds its C ABI with ctypes for tests, the bench
and scripting: every computation runs in the shared library on the GPU.
There is no CPU fallback; importing works without a GPU, calling a compute
function without the built library or without a GPU raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libqgtl_mi355x.so")

GATE = dict(I=0, X=1, Y=2, Z=3, H=4, S=5, T=6, RX=7, RY=8, RZ=9, CNOT=10, CY=11,
            CZ=12, SWAP=13, U1=15, PHASE=19, SDG=26, TDG=27, SX=29)

MODE_MIXED = 0
MODE_FP64_PLAIN = 1

OP_DTYPE = np.dtype([("type", np.int32), ("q0", np.int32), ("q1", np.int32),
                     ("param", np.int32), ("angle", np.float64)], align=True)


class QGTLError(RuntimeError):
    """A negative qgt_error_t return, with the library's message."""

    def __init__(self, code, msg):
        super().__init__(f"qgtl error {code}: {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("sweep_ms", ctypes.c_double),
                ("inject_ms", ctypes.c_double), ("gram_ms", ctypes.c_double),
                ("finalize_ms", ctypes.c_double), ("other_ms", ctypes.c_double),
                ("sweep_launches", ctypes.c_int64), ("sweep_state_passes", ctypes.c_int64),
                ("sweep_bytes", ctypes.c_double), ("gram_launches", ctypes.c_int64),
                ("gram_flops", ctypes.c_double), ("gram_kernel_ms", ctypes.c_double),
                ("kernel_launches", ctypes.c_int64), ("P", ctypes.c_int32), ("n", ctypes.c_int32),
                ("presplit", ctypes.c_int32), ("meet_block", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tune(ctypes.Structure):
    _fields_ = [("chunk_states", ctypes.c_int32), ("gram_chunks", ctypes.c_int32),
                ("gram_variant", ctypes.c_int32), ("sweep_form", ctypes.c_int32), ("no_presplit", ctypes.c_int32),
                ("sweep_schedule", ctypes.c_int32), ("no_fused_inject", ctypes.c_int32),
                ("no_graph", ctypes.c_int32)]


_lib = None


def lib():
    """Load the shared library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libqgtl_mi355x.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.qgtl_init.argtypes = [ip]
        L.qgtl_last_error.restype = ctypes.c_char_p
        L.qgtl_set_profiling.argtypes = [ip]
        L.qgtl_hea_num_ops.argtypes = [ip, ip]
        L.qgtl_hea_ops.argtypes = [ip, ip, vp, vp]
        L.qgtl_num_params.argtypes = [ip, vp]
        L.qgtl_qgt.argtypes = [ip, ip, vp, ip, vp, ctypes.POINTER(Stats)]
        L.qgtl_statevector.argtypes = [ip, ip, vp, ip, vp]
        L.qgtl_plan_create.argtypes = [ip, ip, vp, ip, ctypes.POINTER(Tune)]
        L.qgtl_plan_create.restype = vp
        L.qgtl_plan_set_angles.argtypes = [vp, ip, vp]
        L.qgtl_plan_execute.argtypes = [vp, vp, ctypes.POINTER(Stats)]
        L.qgtl_plan_num_params.argtypes = [vp]
        L.qgtl_plan_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise QGTLError(rc, lib().qgtl_last_error().decode(errors="replace"))


def init(device=0):
    _check(lib().qgtl_init(device))


def set_profiling(enable):
    lib().qgtl_set_profiling(1 if enable else 0)


def hea_ops(n, layers, theta):
    """Hardware-efficient ansatz: RY(q), RZ(q) on every qubit, then CNOT(i -> i+1 mod n)."""
    theta = np.ascontiguousarray(theta, np.float64)
    if theta.size != 2 * n * layers:
        raise ValueError("theta must hold 2*n*layers angles")
    ops = np.zeros(lib().qgtl_hea_num_ops(n, layers), OP_DTYPE)
    lib().qgtl_hea_ops(n, layers, _p(theta), _p(ops))
    return ops


def make_ops(gates):
    """gates: iterable of (name, q0, q1, param, angle)."""
    gates = list(gates)
    ops = np.zeros(len(gates), OP_DTYPE)
    for i, (name, q0, q1, param, angle) in enumerate(gates):
        ops[i] = (GATE[name], q0, q1, param, angle)
    return ops


def num_params(ops):
    return int(np.sum(ops["param"] >= 0))


def qgt(n, ops, mode=MODE_MIXED, stats=False):
    """Full P x P quantum geometric tensor (complex128)."""
    ops = np.ascontiguousarray(ops, OP_DTYPE)
    P = num_params(ops)
    Q = np.zeros((P, P), np.complex128)
    st = Stats()
    _check(lib().qgtl_qgt(n, len(ops), _p(ops), mode, _p(Q), ctypes.byref(st)))
    return (Q, st.as_dict()) if stats else Q


def statevector(n, ops, mode=MODE_MIXED):
    ops = np.ascontiguousarray(ops, OP_DTYPE)
    psi = np.zeros(1 << n, np.complex128)
    _check(lib().qgtl_statevector(n, len(ops), _p(ops), mode, _p(psi)))
    return psi


class Plan:
    """A compiled circuit with device buffers, re-run with new angles."""

    def __init__(self, n, ops, mode=MODE_MIXED, chunk_states=0, gram_chunks=0, gram_terms=0, sweep_form=0,
                 presplit=True, two_sided=True, fused_inject=True, graph=True):
        self.ops = np.ascontiguousarray(ops, OP_DTYPE).copy()
        self.n = n
        tune = Tune(chunk_states, gram_chunks, gram_terms, sweep_form, 0 if presplit else 1, 0 if two_sided else 1,
                    0 if fused_inject else 1, 0 if graph else 1)
        h = lib().qgtl_plan_create(n, len(self.ops), _p(self.ops), mode, ctypes.byref(tune))
        if not h:
            raise QGTLError(-1, lib().qgtl_last_error().decode(errors="replace"))
        self._h = h
        self.P = lib().qgtl_plan_num_params(h)

    def set_angles(self, angles):
        angles = np.ascontiguousarray(angles, np.float64)
        _check(lib().qgtl_plan_set_angles(self._h, len(angles), _p(angles)))

    def execute(self, want_q=True, out=None):
        """Q (P x P complex128, parameter order) and the call's stats; out: a C-contiguous
        complex128 P x P array to write Q into instead of a new one."""
        st = Stats()
        if out is not None:
            if out.dtype != np.complex128 or out.shape != (self.P, self.P) or not out.flags["C_CONTIGUOUS"]:
                raise ValueError("out must be a C-contiguous complex128 array of shape (P, P)")
            Q = out
        else:
            Q = np.empty((self.P, self.P), np.complex128) if want_q else None
        _check(lib().qgtl_plan_execute(self._h, _p(Q) if Q is not None else None, ctypes.byref(st)))
        return Q, st.as_dict()

    def set_hamiltonian(self, terms):
        """terms: array of PAULI_DTYPE (see pauli_terms / tfim_terms)."""
        terms = np.ascontiguousarray(terms, PAULI_DTYPE)
        L = lib()
        L.qgtl_plan_set_hamiltonian.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _check(L.qgtl_plan_set_hamiltonian(self._h, len(terms), _p(terms)))

    def overlaps(self):
        """a_k = <d_k psi|psi> (complex128, parameter order) that the last mixed-mode
        execute() subtracted in Q = G - a a^H."""
        a = np.zeros(2 * self.P, np.float64)
        L = lib()
        L.qgtl_plan_overlaps.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _check(L.qgtl_plan_overlaps(self._h, _p(a)))
        return a.view(np.complex128)

    def energy_gradient(self):
        """E and grad_k = 2 Re <d_k psi|H|psi> from the last execute()."""
        e = np.zeros(1, np.float64)
        g = np.zeros(self.P, np.float64)
        L = lib()
        L.qgtl_plan_energy_gradient.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _check(L.qgtl_plan_energy_gradient(self._h, _p(e), _p(g)))
        return float(e[0]), g

    def close(self):
        if getattr(self, "_h", None):
            lib().qgtl_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PAULI_DTYPE = np.dtype([("xmask", np.uint32), ("zmask", np.uint32), ("coeff_re", np.float64),
                        ("coeff_im", np.float64)], align=True)


def pauli_terms(strings, coeffs):
    """QGTL convention: character 0 of each string acts on the highest qubit."""
    out = np.zeros(len(strings), PAULI_DTYPE)
    L = lib()
    L.qgtl_pauli_term_from_string.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_void_p]
    for i, (s, c) in enumerate(zip(strings, coeffs)):
        _check(L.qgtl_pauli_term_from_string(s.encode(), float(c), out[i:i + 1].ctypes.data))
    return out


def tfim_terms(n, J, h):
    """H = -J sum_{i<n-1} Z_i Z_{i+1} - h sum_i X_i."""
    out = np.zeros(2 * n - 1, PAULI_DTYPE)
    L = lib()
    L.qgtl_tfim_terms.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    k = L.qgtl_tfim_terms(n, J, h, _p(out))
    if k < 0:
        _check(k)
    return out


SOLVE_CHOLESKY, SOLVE_CG, SOLVE_JACOBI_PCG, SOLVE_PCG = 0, 1, 2, 3


class SolveInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("rel_residual", ctypes.c_double), ("device_ms", ctypes.c_double)]


def solve(A, b, lam=0.0, method=SOLVE_CHOLESKY, tol=1e-8, maxit=10000, d=None, U=None, check=True):
    """(A + lam I) x = b on the GPU in FP64; A is (P, P) or (batch, P, P)."""
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    batch = 1 if A.ndim == 2 else A.shape[0]
    P = A.shape[-1]
    x = np.zeros((batch, P), np.float64)
    info = (SolveInfo * batch)()
    dd = None if d is None else np.ascontiguousarray(d, np.float64)
    UU = None if U is None else np.ascontiguousarray(U, np.float64)
    rank = 0 if UU is None else UU.shape[-1]
    L = lib()
    L.qgtl_solve_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rc = L.qgtl_solve_f64(P, batch, _p(A), _p(b), lam, method, tol, maxit,
                          None if dd is None else _p(dd), None if UU is None else _p(UU), rank, _p(x),
                          ctypes.cast(info, ctypes.c_void_p))
    if check:
        _check(rc)
    infos = [dict(status=i.status, iterations=i.iterations, rel_residual=i.rel_residual, device_ms=i.device_ms)
             for i in info]
    return (x[0] if A.ndim == 2 else x), (infos[0] if A.ndim == 2 else infos)


def hermitian_eig(A):
    A = np.ascontiguousarray(A, np.complex128)
    P = A.shape[0]
    ev = np.zeros(P, np.float64)
    V = np.zeros((P, P), np.complex128)
    L = lib()
    L.qgtl_hermitian_eig.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    _check(L.qgtl_hermitian_eig(P, _p(A), _p(ev), _p(V)))
    return ev, V


def rel_err(Q, Qref):
    """Normwise relative error max|Q - Qref| / max|Qref|."""
    return float(np.max(np.abs(Q - Qref)) / np.max(np.abs(Qref)))


# ---------------------------------------------------------------- batches
class BatchTune(ctypes.Structure):
    _fields_ = [("pass_grid", ctypes.c_int32), ("inject_grid", ctypes.c_int32), ("gram_grid", ctypes.c_int32),
                ("gram_chunk_len", ctypes.c_int32), ("pass_variant", ctypes.c_int32), ("no_graph", ctypes.c_int32),
                ("sweep_form", ctypes.c_int32), ("sweep_schedule", ctypes.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("device_ms", ctypes.c_double), ("upload_ms", ctypes.c_double),
                ("download_ms", ctypes.c_double), ("launches", ctypes.c_int64), ("units", ctypes.c_int64),
                ("sweep_bytes", ctypes.c_double), ("gram_flops", ctypes.c_double), ("instances", ctypes.c_int32),
                ("graph", ctypes.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def _batch_lib():
    L = lib()
    if not getattr(L, "_batch_bound", False):
        vp, ip = ctypes.c_void_p, ctypes.c_int
        L.qgtl_batch_create.argtypes = [ip, vp, vp, vp, ctypes.POINTER(BatchTune)]
        L.qgtl_batch_create.restype = vp
        L.qgtl_batch_num_params.argtypes = [vp, ip]
        L.qgtl_batch_q_offset.argtypes = [vp, ip]
        L.qgtl_batch_q_offset.restype = ctypes.c_size_t
        L.qgtl_batch_q_size.argtypes = [vp]
        L.qgtl_batch_q_size.restype = ctypes.c_size_t
        L.qgtl_batch_device_bytes.argtypes = [vp]
        L.qgtl_batch_device_bytes.restype = ctypes.c_size_t
        L.qgtl_batch_set_angles.argtypes = [vp, vp]
        L.qgtl_batch_execute.argtypes = [vp, vp, ctypes.POINTER(BatchStats)]
        L.qgtl_batch_set_tune.argtypes = [vp, ctypes.POINTER(BatchTune)]
        L.qgtl_batch_get_tune.argtypes = [vp, ctypes.POINTER(BatchTune)]
        L.qgtl_batch_autotune.argtypes = [vp, ip, ctypes.c_char_p, ctypes.POINTER(BatchTune)]
        L.qgtl_batch_tune_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(BatchTune)]
        L.qgtl_batch_destroy.argtypes = [vp]
        L._batch_bound = True
    return L


class Batch:
    """Many independent circuits, executed together (grouped launches in one HIP graph).

    circuits: list of (n, ops) with ops from hea_ops / make_ops.
    """

    def __init__(self, circuits, tune=None):
        L = _batch_lib()
        self.n = np.array([c[0] for c in circuits], np.int32)
        self.nops = np.array([len(c[1]) for c in circuits], np.int32)
        self.ops = np.ascontiguousarray(np.concatenate([np.asarray(c[1], OP_DTYPE) for c in circuits]), OP_DTYPE)
        t = tune if tune is not None else BatchTune()
        h = L.qgtl_batch_create(len(circuits), _p(self.n), _p(self.nops), _p(self.ops), ctypes.byref(t))
        if not h:
            raise QGTLError(-1, L.qgtl_last_error().decode(errors="replace"))
        self._h = h
        self.count = len(circuits)
        self.P = [L.qgtl_batch_num_params(h, i) for i in range(self.count)]
        self.q_off = [L.qgtl_batch_q_offset(h, i) for i in range(self.count)]
        self.q_size = L.qgtl_batch_q_size(h)
        self.device_bytes = L.qgtl_batch_device_bytes(h)

    def set_angles(self, per_op_angles):
        """per_op_angles: one array per instance, one angle per op (as in the ops' angle field)."""
        a = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in per_op_angles]))
        if a.size != self.ops.size:
            raise ValueError("one angle per op of every instance")
        _check(_batch_lib().qgtl_batch_set_angles(self._h, _p(a)))

    def execute(self, want_q=True, out=None):
        """Run the batch; returns ([Q_i], stats).  out: a complex128 array of q_size
        elements to write the matrices into (reused across calls instead of a new
        array per call); the returned Q_i are views of it."""
        st = BatchStats()
        if want_q and out is not None:
            if out.dtype != np.complex128 or out.size != self.q_size or not out.flags.c_contiguous:
                raise ValueError("out must be a contiguous complex128 array of q_size elements")
            buf = out.reshape(-1)
        else:
            buf = np.zeros(self.q_size, np.complex128) if want_q else None
        _check(_batch_lib().qgtl_batch_execute(self._h, _p(buf) if want_q else None, ctypes.byref(st)))
        Qs = None
        if want_q:
            Qs = [buf[o:o + P * P].reshape(P, P) for o, P in zip(self.q_off, self.P)]
        return Qs, st.as_dict()

    def autotune(self, reps=3, path=None):
        best = BatchTune()
        _check(_batch_lib().qgtl_batch_autotune(self._h, reps, path.encode() if path else None, ctypes.byref(best)))
        return best

    def set_tune(self, tune):
        _check(_batch_lib().qgtl_batch_set_tune(self._h, ctypes.byref(tune)))

    def get_tune(self):
        t = BatchTune()
        _check(_batch_lib().qgtl_batch_get_tune(self._h, ctypes.byref(t)))
        return t

    def close(self):
        if getattr(self, "_h", None):
            _batch_lib().qgtl_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_batch_tune(path):
    t = BatchTune()
    _check(_batch_lib().qgtl_batch_tune_load(path.encode(), ctypes.byref(t)))
    return t


# ---------------------------------------------------------------- learned preconditioner
def _mlp_lib():
    L = lib()
    if not getattr(L, "_mlp_bound", False):
        vp, ip = ctypes.c_void_p, ctypes.c_int
        L.qgtl_mlp_create.argtypes = [ip, ip, ip, ip, ip, ctypes.c_uint64]
        L.qgtl_mlp_create.restype = vp
        L.qgtl_mlp_num_params.argtypes = [vp]
        L.qgtl_mlp_num_params.restype = ctypes.c_size_t
        L.qgtl_mlp_get_params.argtypes = [vp, vp]
        L.qgtl_mlp_set_params.argtypes = [vp, vp]
        L.qgtl_mlp_set_data.argtypes = [vp, ip, vp, vp]
        L.qgtl_mlp_loss_grad.argtypes = [vp, ip, vp, vp, vp]
        L.qgtl_mlp_train_step.argtypes = [vp, ip, vp, ctypes.c_float, vp]
        L.qgtl_mlp_forward.argtypes = [vp, ip, vp, vp]
        L.qgtl_pc_features.argtypes = [ip, vp, vp, ctypes.c_double, vp]
        L.qgtl_mlp_preconditioner.argtypes = [vp, vp, vp, vp, vp]
        L.qgtl_mlp_destroy.argtypes = [vp]
        L.qgtl_mlp_save.argtypes = [vp, ctypes.c_char_p, ip]
        L.qgtl_mlp_load.argtypes = [ctypes.c_char_p]
        L.qgtl_mlp_load.restype = vp
        L.qgtl_mlp_get_shape.argtypes = [vp, vp, vp, vp, vp, vp]
        L._mlp_bound = True
    return L


def pc_features(theta, D, field):
    """Features of one parameter point: sin theta, cos theta, log diag(A), field (3P + 1 floats)."""
    theta = np.ascontiguousarray(theta, np.float64)
    D = np.ascontiguousarray(D, np.float64)
    x = np.zeros(3 * theta.size + 1, np.float32)
    _check(_mlp_lib().qgtl_pc_features(theta.size, _p(theta), _p(D), float(field), _p(x)))
    return x


class MLP:
    """The learned preconditioner's network, trained and evaluated on the GPU."""

    def __init__(self, in_dim, hidden, hidden_layers, P, rank, seed=0, _handle=None):
        L = _mlp_lib()
        h = _handle if _handle else L.qgtl_mlp_create(in_dim, hidden, hidden_layers, P, rank, seed)
        if not h:
            raise QGTLError(-1, L.qgtl_last_error().decode(errors="replace"))
        self._h = h
        self.in_dim, self.hidden, self.hidden_layers, self.P, self.rank = in_dim, hidden, hidden_layers, P, rank
        self.out_dim = P * (1 + rank)
        self.num_params = L.qgtl_mlp_num_params(h)

    def save(self, path, with_optimizer=True):
        """Checkpoint: fp32 master weights, plus the Adam moments and step when with_optimizer."""
        _check(_mlp_lib().qgtl_mlp_save(self._h, os.fsencode(path), int(bool(with_optimizer))))

    @classmethod
    def load(cls, path):
        """A network of the saved shape with the saved weights (and optimizer state, if saved)."""
        L = _mlp_lib()
        h = L.qgtl_mlp_load(os.fsencode(path))
        if not h:
            raise QGTLError(-1, L.qgtl_last_error().decode(errors="replace"))
        shape = [ctypes.c_int(0) for _ in range(5)]
        L.qgtl_mlp_get_shape(h, *[ctypes.byref(v) for v in shape])
        return cls(*[v.value for v in shape], _handle=h)

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        _check(_mlp_lib().qgtl_mlp_get_params(self._h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params:
            raise ValueError("parameter count mismatch")
        _check(_mlp_lib().qgtl_mlp_set_params(self._h, _p(p)))

    def set_data(self, X, As):
        X = np.ascontiguousarray(X, np.float32)
        As = np.ascontiguousarray(As, np.float32)
        if X.shape[1] != self.in_dim or As.shape[1:] != (self.P, self.P) or X.shape[0] != As.shape[0]:
            raise ValueError("X must be S x in_dim and As S x P x P")
        _check(_mlp_lib().qgtl_mlp_set_data(self._h, X.shape[0], _p(X), _p(As)))
        self.S = X.shape[0]

    def loss_grad(self, idx):
        idx = np.ascontiguousarray(idx, np.int32)
        loss = np.zeros(1)
        g = np.zeros(self.num_params, np.float32)
        _check(_mlp_lib().qgtl_mlp_loss_grad(self._h, idx.size, _p(idx), _p(loss), _p(g)))
        return float(loss[0]), g

    def train_step(self, idx, lr):
        idx = np.ascontiguousarray(idx, np.int32)
        loss = np.zeros(1)
        _check(_mlp_lib().qgtl_mlp_train_step(self._h, idx.size, _p(idx), float(lr), _p(loss)))
        return float(loss[0])

    def forward(self, X):
        X = np.ascontiguousarray(np.atleast_2d(X), np.float32)
        Y = np.zeros((X.shape[0], self.out_dim), np.float32)
        _check(_mlp_lib().qgtl_mlp_forward(self._h, X.shape[0], _p(X), _p(Y)))
        return Y

    def preconditioner(self, x, D):
        x = np.ascontiguousarray(x, np.float32)
        D = np.ascontiguousarray(D, np.float64)
        d = np.zeros(self.P)
        U = np.zeros((self.P, self.rank))
        _check(_mlp_lib().qgtl_mlp_preconditioner(self._h, _p(x), _p(D), _p(d), _p(U)))
        return d, U

    def close(self):
        if getattr(self, "_h", None):
            _mlp_lib().qgtl_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- files
def _io_lib():
    L = lib()
    if not getattr(L, "_io_bound", False):
        vp, ip, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p
        ipp = ctypes.POINTER(ctypes.c_int)
        L.qgtl_qgt_save.argtypes = [cp, ip, ip, vp, vp]
        L.qgtl_qgt_file_info.argtypes = [cp, ipp, ipp, ipp]
        L.qgtl_qgt_load.argtypes = [cp, ip, vp, vp]
        L.qgtl_dataset_save.argtypes = [cp, ip, ip, ip, vp, vp]
        L.qgtl_dataset_file_info.argtypes = [cp, ipp, ipp, ipp]
        L.qgtl_dataset_load.argtypes = [cp, ip, ip, ip, vp, vp]
        L._io_bound = True
    return L


def qgt_save(path, n, Q, theta=None):
    """Write a P x P complex QGT (and optionally its angles) in the library's QGT file format."""
    Q = np.ascontiguousarray(Q, np.complex128)
    th = None if theta is None else np.ascontiguousarray(theta, np.float64)
    _check(_io_lib().qgtl_qgt_save(os.fsencode(path), int(n), Q.shape[0], None if th is None else _p(th), _p(Q)))


def qgt_load(path):
    """(n, Q, theta or None) from a QGT file; size, format and checksum are checked."""
    L = _io_lib()
    n, P, ht = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _check(L.qgtl_qgt_file_info(os.fsencode(path), ctypes.byref(n), ctypes.byref(P), ctypes.byref(ht)))
    Q = np.zeros((P.value, P.value), np.complex128)
    th = np.zeros(P.value) if ht.value else None
    _check(L.qgtl_qgt_load(os.fsencode(path), P.value, None if th is None else _p(th), _p(Q)))
    return n.value, Q, th


def dataset_save(path, X, As):
    """Training pairs of the learned preconditioner: features X (S x F) and scaled matrices As (S x P x P)."""
    X = np.ascontiguousarray(X, np.float32)
    As = np.ascontiguousarray(As, np.float32)
    if X.ndim != 2 or As.ndim != 3 or As.shape[0] != X.shape[0] or As.shape[1] != As.shape[2]:
        raise ValueError("X must be S x F and As S x P x P")
    _check(_io_lib().qgtl_dataset_save(os.fsencode(path), X.shape[0], X.shape[1], As.shape[1], _p(X), _p(As)))


def dataset_load(path):
    L = _io_lib()
    S, F, P = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _check(L.qgtl_dataset_file_info(os.fsencode(path), ctypes.byref(S), ctypes.byref(F), ctypes.byref(P)))
    X = np.zeros((S.value, F.value), np.float32)
    As = np.zeros((S.value, P.value, P.value), np.float32)
    _check(L.qgtl_dataset_load(os.fsencode(path), S.value, F.value, P.value, _p(X), _p(As)))
    return X, As


# ---------------------------------------------------------------- natural-gradient VQE step
class NGConfig(ctypes.Structure):
    _fields_ = [("lam", ctypes.c_double), ("eta", ctypes.c_double), ("method", ctypes.c_int32),
                ("maxit", ctypes.c_int32), ("tol", ctypes.c_double)]


class NGInfo(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("grad_norm", ctypes.c_double), ("step_norm", ctypes.c_double),
                ("solve_iterations", ctypes.c_int32), ("solve_status", ctypes.c_int32),
                ("fallback", ctypes.c_int32), ("pad", ctypes.c_int32), ("solve_rel_residual", ctypes.c_double), ("solve_device_ms", ctypes.c_double),
                ("qgt_ms", ctypes.c_double), ("solve_ms", ctypes.c_double), ("step_ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def ng_step(plan, theta, lam=1e-3, eta=0.1, method=0, tol=1e-8, maxit=10000, mlp=None, field=0.0):
    """One natural-gradient step on `plan` (Hamiltonian set): updates theta in place, returns NGInfo dict."""
    L = lib()
    if not getattr(L, "_ng_bound", False):
        L.qgtl_ng_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(NGConfig), ctypes.c_void_p,
                                   ctypes.c_double, ctypes.POINTER(NGInfo)]
        L._ng_bound = True
    if theta.dtype != np.float64 or not theta.flags["C_CONTIGUOUS"]:
        raise ValueError("theta must be a C-contiguous float64 array (updated in place)")
    cfg = NGConfig(lam, eta, method, maxit, tol)
    info = NGInfo()
    _check(L.qgtl_ng_step(plan._h, _p(theta), ctypes.byref(cfg), mlp._h if mlp is not None else None, float(field),
                          ctypes.byref(info)))
    return info.as_dict()
