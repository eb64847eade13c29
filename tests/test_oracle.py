"""Pins of the FP64 CPU oracle (no GPU): closed forms, identities, brute force,
QGTL's own state-magnitude tests, and the committed golden fixtures.

QGTL's tests pin no QGT value (tests/test_quantum_geometric_minimal.c only
prints one), so the QGT is pinned here by mathematics: each check below would
fail on a dropped term, a wrong sign, a wrong qubit index or a transposed
operand.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

X = np.array([[0, 1], [1, 0]], complex)
Y = np.array([[0, -1j], [1j, 0]])
Z = np.diag([1.0 + 0j, -1.0])
I2 = np.eye(2)


def test_single_qubit_closed_form():
    # RY(a) then RZ(b) on |0>: Q = [[1/4, i sin(a)/4], [-i sin(a)/4, sin^2(a)/4]]
    for a, b in [(0.7, 1.3), (2.9, -0.4), (np.pi / 2, 0.0)]:
        ops = O.make_ops([("RY", 0, -1, 0, a), ("RZ", 0, -1, 1, b)])
        Q = O.qgt(1, ops)
        expect = np.array([[0.25, 0.25j * np.sin(a)], [-0.25j * np.sin(a), np.sin(a) ** 2 / 4]])
        np.testing.assert_allclose(Q, expect, atol=1e-15)


@pytest.mark.parametrize("n", [3, 7, 12])
def test_layer1_product_block(n):
    """Q_mn depends only on the circuit up to max(m, n); before the first CNOT
    ring the state is a product state, so the leading 2n x 2n block is the
    per-qubit single-qubit closed form and zero across qubits."""
    L = 2
    th = O.uniform_angles(11 + n, 2 * n * L)
    Q = O.qgt(n, O.hea_ops(n, L, th))
    B = Q[: 2 * n, : 2 * n]
    expect = np.zeros((2 * n, 2 * n), complex)
    for q in range(n):
        a = th[q]
        expect[q, q] = 0.25
        expect[q, n + q] = 0.25j * np.sin(a)
        expect[n + q, q] = -0.25j * np.sin(a)
        expect[n + q, n + q] = np.sin(a) ** 2 / 4
    np.testing.assert_allclose(B, expect, atol=1e-14)


def test_shift_identity_and_finite_differences():
    n, L = 5, 3
    th = O.uniform_angles(3, 2 * n * L)
    ops = O.hea_ops(n, L, th)
    psi, D = O.derivative_states(n, ops)
    Ds = O.derivative_states_shift(n, ops)
    np.testing.assert_allclose(D, Ds, atol=1e-14)
    h = 1e-5
    for k in [0, 7, 2 * n * L - 1]:
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (O.run(n, O.hea_ops(n, L, tp)) - O.run(n, O.hea_ops(n, L, tm))) / (2 * h)
        np.testing.assert_allclose(D[k], fd, atol=1e-9)
    np.testing.assert_allclose(psi, O.run(n, ops), atol=1e-15)


def _embed(n, q, g):
    return np.kron(np.kron(np.eye(1 << (n - 1 - q)), g), np.eye(1 << q))


def _dense_gate(n, name, q0, q1, angle):
    c, s = np.cos(angle / 2), np.sin(angle / 2)
    one = {"RX": np.array([[c, -1j * s], [-1j * s, c]]), "RY": np.array([[c, -s], [s, c]], complex),
           "RZ": np.diag([np.exp(-0.5j * angle), np.exp(0.5j * angle)]), "H": np.array([[1, 1], [1, -1]]) / np.sqrt(2),
           "X": X, "Y": Y, "Z": Z, "S": np.diag([1, 1j]), "T": np.diag([1, np.exp(0.25j * np.pi)])}
    if name in one:
        return _embed(n, q0, one[name])
    P0, P1 = np.diag([1.0, 0.0]), np.diag([0.0, 1.0])
    if name == "CNOT":
        return _embed(n, q0, P0) + _embed(n, q0, P1) @ _embed(n, q1, X)
    if name == "CZ":
        return _embed(n, q0, P0) + _embed(n, q0, P1) @ _embed(n, q1, Z)
    raise ValueError(name)


@pytest.mark.parametrize("seed", [0, 1])
def test_dense_bruteforce_qgt(seed):
    """Independent numpy brute force (Kronecker products, generator insertion)
    for a mixed circuit on 4 qubits."""
    rng = np.random.default_rng(seed)
    n = 4
    gates = [("H", 0, -1, -1, 0.0), ("RY", 1, -1, 0, rng.uniform(0, 6)), ("RX", 2, -1, 1, rng.uniform(0, 6)),
             ("CNOT", 0, 2, -1, 0.0), ("RZ", 0, -1, 2, rng.uniform(0, 6)), ("CZ", 3, 1, -1, 0.0),
             ("RY", 3, -1, 3, rng.uniform(0, 6)), ("T", 1, -1, -1, 0.0), ("RX", 1, -1, 4, rng.uniform(0, 6)),
             ("CNOT", 3, 0, -1, 0.0), ("RZ", 2, -1, 5, rng.uniform(0, 6)), ("S", 3, -1, -1, 0.0),
             ("RY", 0, -1, 6, rng.uniform(0, 6))]
    ops = O.make_ops(gates)
    mats = [_dense_gate(n, g[0], g[1], g[2], g[4]) for g in gates]
    gen = {"RX": X, "RY": Y, "RZ": Z}
    e0 = np.zeros(1 << n, complex)
    e0[0] = 1
    psi = e0.copy()
    for M in mats:
        psi = M @ psi
    D = []
    for gi, g in enumerate(gates):
        if g[3] < 0:
            continue
        v = e0.copy()
        for M in mats[: gi + 1]:
            v = M @ v
        v = -0.5j * (_embed(n, g[1], gen[g[0]]) @ v)
        for M in mats[gi + 1:]:
            v = M @ v
        D.append(v)
    D = np.array(D)
    Qbf = D.conj() @ D.T - np.outer(D.conj() @ psi, psi.conj() @ D.T)
    np.testing.assert_allclose(O.qgt(n, ops), Qbf, atol=1e-14)
    np.testing.assert_allclose(O.dense_unitary(n, ops)[:, 0], psi, atol=1e-14)


def test_qgt_invariants_and_variants():
    n, L = 8, 3
    th = O.uniform_angles(5, 2 * n * L)
    ops = O.hea_ops(n, L, th)
    psi, D = O.derivative_states(n, ops)
    Q = O.qgt(n, ops)
    np.testing.assert_allclose(Q, Q.conj().T, atol=1e-15)
    assert np.linalg.eigvalsh(Q).min() > -1e-13
    d = np.real(np.diag(Q))
    assert d.min() > -1e-15 and d.max() <= 0.25 + 1e-15
    np.testing.assert_allclose(np.sum(np.abs(D) ** 2, axis=1), 0.25, atol=1e-14)
    assert np.max(np.abs(np.real(D.conj() @ psi))) < 1e-14          # Re<d|psi> = 0
    np.testing.assert_allclose(O.qgt_from_states(psi, D, "pairwise"), Q, atol=1e-15)
    np.testing.assert_allclose(O.qgt_hoisted(n, ops), Q, atol=1e-15)
    rows, cols = [1, 9, 30], [0, 4, 47]
    np.testing.assert_allclose(O.qgt_block(n, ops, rows, cols), Q[np.ix_(rows, cols)], atol=1e-15)
    # global phase invariance: prepend Z then RZ on qubit 0 changes only the phase of |0..0>
    ops2 = O.make_ops([("RZ", 0, -1, -1, 1.234)])
    ops2 = np.concatenate([ops2, ops])
    np.testing.assert_allclose(O.qgt(n, ops2), Q, atol=1e-14)


def test_reference_state_magnitudes():
    # tests/test_quantum_simulator_cpu.c:126-134 (Bell) and :404-407 (GHZ)
    bell = O.run(2, O.make_ops([("H", 0, -1, -1, 0.0), ("CNOT", 0, 1, -1, 0.0)]))
    np.testing.assert_allclose(np.abs(bell), [2 ** -0.5, 0, 0, 2 ** -0.5], atol=1e-10)
    ghz = O.run(3, O.make_ops([("H", 0, -1, -1, 0.0), ("CNOT", 0, 1, -1, 0.0), ("CNOT", 1, 2, -1, 0.0)]))
    expect = np.zeros(8)
    expect[0] = expect[7] = 2 ** -0.5
    np.testing.assert_allclose(np.abs(ghz), expect, atol=1e-10)
    # tests/test_quantum_gates.c:26-79: X and H on |0>
    np.testing.assert_allclose(np.abs(O.run(1, O.make_ops([("X", 0, -1, -1, 0.0)]))), [0, 1], atol=1e-12)


def test_ring_index_math():
    """The sweep kernel stores the CNOT ring through a closed form: for a tile
    over bits 0..m-1 and tile index r, output = (y_hi << m) | y_lo with
    y_lo = local map, y_hi = prefix_xor(r) complemented when y_{m-1} = 1.
    Restated here and checked against the gate-by-gate ring on bit vectors."""
    for n, m in [(5, 3), (9, 4), (14, 13), (20, 13)]:
        N = 1 << n
        x = np.arange(N, dtype=np.int64)
        y = x.copy()
        for i in range(n):
            y ^= ((y >> i) & 1) << ((i + 1) % n)
        r = x >> m
        pre = r.copy()
        for s in (1, 2, 4, 8, 16):
            pre ^= pre << s
        nh = n - m
        ones = (1 << nh) - 1
        pre &= ones
        par_r = np.zeros_like(r)
        t = r.copy()
        while t.any():
            par_r ^= t & 1
            t >>= 1
        xl = x & ((1 << m) - 1)
        p = xl.copy()
        for s in (1, 2, 4, 8):
            p ^= p << s
        p &= (1 << m) - 1
        ylo = p ^ (((p >> (m - 1)) & 1) ^ par_r)     # bit 0 flipped by P_{m-1} ^ par(r)
        yhi = np.where((ylo >> (m - 1)) & 1, pre ^ ones, pre)
        np.testing.assert_array_equal((yhi << m) | ylo, y)
        # the kernel's inverse of the local map
        x0 = (ylo & 1) ^ ((ylo >> (m - 1)) & 1) ^ par_r
        x1 = ((ylo >> 1) & 1) ^ x0
        rest = (ylo ^ (ylo << 1)) & ~3 & ((1 << m) - 1)
        np.testing.assert_array_equal(rest | (x1 << 1) | x0, xl)


def test_tfim_energy_gradient():
    n, L, J, h = 6, 3, 1.0, 0.8
    th = O.uniform_angles(9, 2 * n * L)
    ops = O.hea_ops(n, L, th)
    psi = O.run(n, ops)
    # energy from QGTL-style Pauli strings (char 0 = highest qubit)
    E = 0.0
    for i in range(n - 1):
        s = ["I"] * n
        s[n - 1 - i] = s[n - 2 - i] = "Z"
        E -= J * O.pauli_expectation(n, psi, "".join(s))
    for i in range(n):
        s = ["I"] * n
        s[n - 1 - i] = "X"
        E -= h * O.pauli_expectation(n, psi, "".join(s))
    e2, g = O.tfim_gradient(n, ops, J, h)
    assert abs(E - e2) < 1e-13 and abs(E - O.tfim_energy(n, J, h, psi)) < 1e-13
    for k in [0, 5, 2 * n * L - 1]:
        tp, tm = th.copy(), th.copy()
        tp[k] += 1e-6
        tm[k] -= 1e-6
        fd = (O.tfim_energy(n, J, h, O.run(n, O.hea_ops(n, L, tp))) -
              O.tfim_energy(n, J, h, O.run(n, O.hea_ops(n, L, tm)))) / 2e-6
        assert abs(fd - g[k]) < 1e-8


def test_solvers():
    rng = np.random.default_rng(0)
    n, L = 6, 3
    Q = O.qgt(n, O.hea_ops(n, L, O.uniform_angles(2, 2 * n * L)))
    A = Q.real + 1e-3 * np.eye(len(Q))
    b = rng.normal(size=len(Q))
    x = O.cholesky_solve(A, b)
    np.testing.assert_allclose(A @ x, b, atol=1e-10)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-12
    xc, it, res = O.pcg(A, b, 1e-8, 100000)
    assert res <= 1e-8 and np.linalg.norm(xc - x) / np.linalg.norm(x) < 1e-5
    xj, itj, resj = O.pcg(A, b, 1e-8, 100000, d=1.0 / np.diag(A))
    assert resj <= 1e-8
    assert O.cholesky_solve(-np.eye(3), np.ones(3)) is None        # not PD
    ev, V = O.hermitian_eig(Q)
    np.testing.assert_allclose(ev, np.linalg.eigvalsh(Q), atol=1e-13)
    np.testing.assert_allclose(V @ np.diag(ev) @ V.conj().T, Q, atol=1e-13)
    g = rng.normal(size=len(Q)) + 0j
    xg, lam = O.regularized_natural_gradient(Q, g, lam=1e-4, adaptive=False)
    np.testing.assert_allclose(xg, np.linalg.solve(Q + 1e-4 * np.eye(len(Q)), g), rtol=1e-8, atol=1e-8)
    assert lam == pytest.approx(1e-4)


def test_golden_fixtures():
    path = os.path.join(GOLDEN, "cfg1_hea_n10_L4.npz")
    f = np.load(path, allow_pickle=False)
    n, L = int(f["n"]), int(f["layers"])
    th = O.uniform_angles(0, 2 * n * L)
    np.testing.assert_array_equal(th, f["theta"])
    Q = O.qgt(n, O.hea_ops(n, L, th))
    np.testing.assert_allclose(Q, f["Q"], atol=1e-14, rtol=0)
    np.testing.assert_allclose(Q.real, f["metric"], atol=1e-14)
    np.testing.assert_allclose(Q.imag, f["berry"], atol=1e-14)
