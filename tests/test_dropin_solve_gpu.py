"""GPU tests of the drop-in C API and the FP64 natural-gradient solvers.

Drop-in: tests/c/test_dropin_api (QGTL's API called from C, as QGTL's own
tests do) must reproduce the FP64 oracle's QGT to 1e-6 normwise (its output is
ComplexFloat), the oracle's state to 1e-5, and its natural-gradient solve.
Solvers: Cholesky to 1e-12 relative residual, CG/PCG to the requested 1e-8,
Hermitian eigenvalues to 1e-10 of numpy/oracle.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import quantum_geometric_tensor_amd as Q

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "c", "_build", "test_dropin_api")


@pytest.mark.parametrize("n,L,seed", [(4, 2, 5), (10, 4, 0), (14, 2, 3)])
def test_dropin_c_api(tmp_path, n, L, seed):
    out = tmp_path / "out.bin"
    res = subprocess.run([PROG, str(n), str(L), str(seed), str(out)], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=120)
    assert res.returncode == 0, res.stdout
    data = np.fromfile(out, np.float64)
    P = int(data[0])
    assert P == 2 * n * L
    off = 1
    Qc = data[off:off + 2 * P * P].view(np.complex128).reshape(P, P)
    off += 2 * P * P
    psi = data[off:off + 2 * (1 << n)].view(np.complex128)
    off += 2 * (1 << n)
    x = data[off:off + 2 * P].view(np.complex128)
    th = O.uniform_angles(seed, P)
    ops = O.hea_ops(n, L, th)
    Qref = O.qgt(n, ops)
    assert Q.rel_err(Qc, Qref) <= 1e-6
    assert np.max(np.abs(psi - O.run(n, ops))) < 1e-5
    # natural gradient on the ComplexFloat metric actually passed: Re Q as float32
    G = Qc.real.astype(np.float32).astype(np.float64)
    g = 1.0 + np.arange(P) / P
    xref = np.linalg.solve(G + 1e-3 * np.eye(P), g)
    assert np.max(np.abs(x.real - xref)) / np.max(np.abs(xref)) < 1e-5
    assert np.max(np.abs(x.imag)) < 1e-6 * np.max(np.abs(xref))


def metric_system(n=8, L=3, lam=1e-3, seed=2):
    th = O.uniform_angles(seed, 2 * n * L)
    Qo = O.qgt(n, O.hea_ops(n, L, th))
    A = Qo.real
    b = np.random.default_rng(seed).normal(size=A.shape[0])
    return A, b, lam


def test_cholesky_and_cg_family():
    A, b, lam = metric_system()
    xo = O.cholesky_solve(A + lam * np.eye(len(b)), b)
    x, info = Q.solve(A, b, lam, Q.SOLVE_CHOLESKY)
    assert info["status"] == 0
    Ar = A + lam * np.eye(len(b))
    assert np.linalg.norm(Ar @ x - b) / np.linalg.norm(b) < 1e-12
    np.testing.assert_allclose(x, xo, rtol=1e-9, atol=1e-9 * np.max(np.abs(xo)))
    for method in (Q.SOLVE_CG, Q.SOLVE_JACOBI_PCG):
        xc, info = Q.solve(A, b, lam, method, tol=1e-8, maxit=20000)
        assert info["rel_residual"] <= 1e-8
        assert np.linalg.norm(Ar @ xc - b) / np.linalg.norm(b) < 2e-8
        xoc, ito, _ = O.pcg(Ar, b, 1e-8, 20000, d=(1.0 / np.diag(Ar)) if method == Q.SOLVE_JACOBI_PCG else None)
        assert abs(info["iterations"] - ito) <= max(3, ito // 20)
    # diag + low-rank preconditioner (deflation of the 8 smallest eigenpairs)
    P = len(b)
    w, V = np.linalg.eigh(Ar)
    U = V[:, :8] / np.sqrt(w[:8])[None, :]
    d = np.full(P, 1.0 / np.diag(Ar).mean())
    xp, info = Q.solve(A, b, lam, Q.SOLVE_PCG, tol=1e-8, maxit=20000, d=d, U=U)
    assert info["rel_residual"] <= 1e-8
    xop, itop, _ = O.pcg(Ar, b, 1e-8, 20000, d=d, U=U)
    assert abs(info["iterations"] - itop) <= max(3, itop // 20)
    # batched: two different systems in one launch
    A2 = np.stack([A, A.T * 0.5 + 0.5 * A])
    b2 = np.stack([b, b[::-1].copy()])
    x2, infos = Q.solve(A2, b2, lam, Q.SOLVE_CHOLESKY)
    for s in range(2):
        np.testing.assert_allclose((A2[s] + lam * np.eye(P)) @ x2[s], b2[s], atol=1e-9)
    # not positive definite -> error with QGTL's code
    with pytest.raises(Q.QGTLError) as e:
        Q.solve(-np.eye(4), np.ones(4), 0.0, Q.SOLVE_CHOLESKY)
    assert e.value.code == -55


@pytest.mark.parametrize("P,rank", [(7, 0), (64, 8), (160, 8), (192, 8), (240, 16), (500, 4), (1024, 8)])
def test_pcg_sizes_and_cache_tiers(P, rank):
    """PCG at sizes where the kernel's A columns sit in registers only (P=7, 64), spill into its
    LDS cache (160) and into global reads (192, 240, 500, 1024), with U in LDS (P*rank <= 2048)
    or read from global memory (240 x 16, 1024 x 8).  Same iteration count as the oracle's PCG
    (within 5%), 1e-8 relative residual, and the true residual checked on the host."""
    rng = np.random.default_rng(P + rank)
    Qm, _ = np.linalg.qr(rng.normal(size=(P, P)))
    ev = np.logspace(-2, 1, P)
    A = (Qm * ev) @ Qm.T
    A = 0.5 * (A + A.T)
    b = rng.normal(size=P)
    d = 1.0 / np.diag(A)
    U = None
    if rank:
        U = Qm[:, :rank] * np.sqrt(1.0 / ev[:rank] - d.mean())[None, :]
    for method, dd, UU in [(Q.SOLVE_JACOBI_PCG, None, None), (Q.SOLVE_PCG, d, U)]:
        if method == Q.SOLVE_PCG and U is None:
            continue
        x, info = Q.solve(A, b, 0.0, method, tol=1e-8, maxit=20000, d=dd, U=UU)
        assert info["rel_residual"] <= 1e-8
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 2e-8
        _, ito, _ = O.pcg(A, b, 1e-8, 20000, d=d if dd is None else dd, U=UU)
        assert abs(info["iterations"] - ito) <= max(3, ito // 20), (P, rank, method, info["iterations"], ito)


def test_hermitian_eig_matches_oracle():
    n, L = 6, 3
    Qo = O.qgt(n, O.hea_ops(n, L, O.uniform_angles(4, 2 * n * L)))
    ev, V = Q.hermitian_eig(Qo)
    np.testing.assert_allclose(ev, np.linalg.eigvalsh(Qo), atol=1e-12)
    np.testing.assert_allclose(V @ np.diag(ev) @ V.conj().T, Qo, atol=1e-12)
    np.testing.assert_allclose(V.conj().T @ V, np.eye(len(ev)), atol=1e-12)


class CF(ctypes.Structure):
    _fields_ = [("real", ctypes.c_float), ("imag", ctypes.c_float)]


class NGConfig(ctypes.Structure):
    _fields_ = [("regularization_param", ctypes.c_float), ("condition_threshold", ctypes.c_float),
                ("use_adaptive_regularization", ctypes.c_bool), ("use_pseudoinverse_fallback", ctypes.c_bool),
                ("singular_value_cutoff", ctypes.c_float)]


def _cf(a):
    a = np.ascontiguousarray(a, np.complex64)
    return a


def test_reference_natural_gradient_api():
    L = Q.lib()
    n, Ln = 5, 2
    Qo = O.qgt(n, O.hea_ops(n, Ln, O.uniform_angles(6, 2 * n * Ln)))
    P = len(Qo)
    M = _cf(Qo)
    g = _cf(np.linspace(1, 2, P) + 0.3j)
    x = np.zeros(P, np.complex64)
    cfg = NGConfig(1e-3, 1e8, False, True, 1e-10)
    L.compute_regularized_natural_gradient.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_size_t, ctypes.POINTER(NGConfig)]
    L.compute_regularized_natural_gradient.restype = ctypes.c_bool
    assert L.compute_regularized_natural_gradient(g.ctypes.data, M.ctypes.data, x.ctypes.data, P, ctypes.byref(cfg))
    Mf = M.astype(np.complex128)
    xref, lam = O.regularized_natural_gradient(Mf, g.astype(np.complex128), lam=1e-3, adaptive=False)
    assert np.max(np.abs(x - xref)) / np.max(np.abs(xref)) < 1e-5
    # condition number, adaptive lambda and pseudo-inverse fallback follow the oracle
    kappa = ctypes.c_float()
    L.compute_complex_matrix_condition_number.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_float)]
    L.compute_complex_matrix_condition_number.restype = ctypes.c_bool
    assert L.compute_complex_matrix_condition_number(M.ctypes.data, P, ctypes.byref(kappa))
    sv = np.abs(np.linalg.eigvalsh(Mf))
    k_ref = sv.max() / sv[sv > 0].min()
    assert abs(kappa.value - k_ref) / k_ref < 1e-3
    cfg2 = NGConfig(1e-4, 1e3, True, True, 1e-10)
    x2 = np.zeros(P, np.complex64)
    assert L.compute_regularized_natural_gradient(g.ctypes.data, M.ctypes.data, x2.ctypes.data, P, ctypes.byref(cfg2))
    x2ref, lam2 = O.regularized_natural_gradient(Mf, g.astype(np.complex128), lam=1e-4, cond=1e3, adaptive=True)
    assert lam2 > 1e-4
    assert np.max(np.abs(x2 - x2ref)) / np.max(np.abs(x2ref)) < 1e-4
    # singular metric without regularisation: LU fails, pseudo-inverse answers
    Z = np.zeros((4, 4), np.complex64)
    Z[0, 0] = 1.0
    Z[1, 1] = 2.0
    gz = _cf(np.ones(4))
    xz = np.zeros(4, np.complex64)
    cfg3 = NGConfig(0.0, 1e8, False, True, 1e-6)
    assert L.compute_regularized_natural_gradient(gz.ctypes.data, Z.ctypes.data, xz.ctypes.data, 4, ctypes.byref(cfg3))
    np.testing.assert_allclose(xz, [1.0, 0.5, 0, 0], atol=1e-6)
    cfg4 = NGConfig(0.0, 1e8, False, False, 1e-6)
    assert not L.compute_regularized_natural_gradient(gz.ctypes.data, Z.ctypes.data, xz.ctypes.data, 4, ctypes.byref(cfg4))
