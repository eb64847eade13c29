"""GPU parity of the HIP QGT path against the FP64 CPU oracle.

Bars (normwise, max|Q - Qref| / max|Qref|): mixed precision <= 1e-6; plain FP64
GPU <= 1e-11.  State vectors: mixed <= 1e-5 absolute, FP64 <= 1e-12.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import quantum_geometric_tensor_amd as Q

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6
FP64_TOL = 1e-11


def hea(n, L, seed=0):
    th = O.uniform_angles(seed, 2 * n * L)
    return th, Q.hea_ops(n, L, th), O.hea_ops(n, L, th)


def test_mfma_lane_layout():
    out = np.zeros(32 * 32, np.float32)
    rc = Q.lib().qgtl_debug_mfma_layout(out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, Q.lib().qgtl_last_error()
    i = np.arange(32)[:, None]
    j = np.arange(32)[None, :]
    expect = (1 + i // 16) * (i % 16 + 16 * (j % 8))
    np.testing.assert_array_equal(out.reshape(32, 32), expect.astype(np.float32))


def test_mfma16_lane_layout():
    out = np.zeros(16 * 16, np.float32)
    rc = Q.lib().qgtl_debug_mfma_layout16(out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, Q.lib().qgtl_last_error()
    i = np.arange(16)[:, None]
    j = np.arange(16)[None, :]
    expect = i + 16 * (i & 1) + 32 * j
    np.testing.assert_array_equal(out.reshape(16, 16), expect.astype(np.float32))


@pytest.mark.parametrize("n,L", [(1, 2), (2, 2), (4, 3), (10, 4), (13, 2), (14, 2), (16, 2)])
def test_statevector_hea(n, L):
    th, ops, oops = hea(n, L, seed=n)
    ref = O.run(n, oops)
    psi64 = Q.statevector(n, ops, Q.MODE_FP64_PLAIN)
    np.testing.assert_allclose(psi64, ref, atol=1e-12, rtol=0)
    psi32 = Q.statevector(n, ops, Q.MODE_MIXED)
    assert np.max(np.abs(psi32 - ref)) < 1e-5


@pytest.mark.parametrize("n,L", [(1, 1), (2, 2), (4, 3), (10, 4)])
def test_qgt_plain_fp64(n, L):
    th, ops, oops = hea(n, L, seed=100 + n)
    ref = O.qgt(n, oops)
    got = Q.qgt(n, ops, Q.MODE_FP64_PLAIN)
    assert Q.rel_err(got, ref) <= FP64_TOL


@pytest.mark.parametrize("n,L", [(1, 1), (2, 2), (3, 3), (4, 3), (8, 3), (10, 4), (13, 2), (14, 2), (15, 2)])
def test_qgt_mixed(n, L):
    th, ops, oops = hea(n, L, seed=200 + n)
    ref = O.qgt(n, oops)
    got = Q.qgt(n, ops, Q.MODE_MIXED)
    err = Q.rel_err(got, ref)
    assert err <= MIXED_TOL, err


# 129 .. 256 Gram states (P + 1 with the primal): one workgroup per split-K chunk computes the
# three tile pairs; P = 128 puts one state past the first tile, P = 252 fills the second
@pytest.mark.parametrize("n,L", [(16, 4), (14, 9), (12, 10)])
def test_qgt_mixed_two_tile_gram(n, L):
    th, ops, oops = hea(n, L, seed=300 + n)
    ref = O.qgt(n, oops)
    got = Q.qgt(n, ops, Q.MODE_MIXED)
    err = Q.rel_err(got, ref)
    print(f"n={n} L={L} P={2 * n * L}: rel err {err:.2e}")
    assert err <= MIXED_TOL, err


def test_qgt_mixed_deterministic():
    th, ops, _ = hea(14, 3, seed=7)
    a = Q.qgt(14, ops, Q.MODE_MIXED)
    b = Q.qgt(14, ops, Q.MODE_MIXED)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,L", [(6, 3), (12, 3), (16, 2)])
def test_energy_gradient_tfim(n, L):
    J, h = 1.0, 0.83
    th, ops, oops = hea(n, L, seed=500 + n)
    e_ref, g_ref = O.tfim_gradient(n, oops, J, h)
    p = Q.Plan(n, ops, Q.MODE_MIXED)
    p.set_hamiltonian(Q.tfim_terms(n, J, h))
    p.execute()
    e, g = p.energy_gradient()
    p.close()
    assert abs(e - e_ref) <= 1e-6 * max(1.0, abs(e_ref))
    assert np.max(np.abs(g - g_ref)) <= 1e-6 * max(1.0, np.max(np.abs(g_ref)))


def test_pauli_strings_qgtl_convention():
    # a general Pauli sum with Y terms, QGTL string convention (char 0 = highest qubit)
    n, L = 5, 2
    th, ops, oops = hea(n, L, seed=77)
    strings, coeffs = ["XYZIZ", "ZZIII", "IYIXI", "YIIIY"], [0.7, -1.1, 0.4, 0.25]
    psi = O.run(n, oops)
    e_ref = sum(c * O.pauli_expectation(n, psi, s) for s, c in zip(strings, coeffs))
    p = Q.Plan(n, ops, Q.MODE_MIXED)
    p.set_hamiltonian(Q.pauli_terms(strings, coeffs))
    p.execute()
    e, g = p.energy_gradient()
    p.close()
    assert abs(e - e_ref) < 1e-6
    # gradient by central differences of the oracle energy
    for k in [0, 7, 19]:
        tp, tm = th.copy(), th.copy()
        tp[k] += 1e-5
        tm[k] -= 1e-5
        ep = sum(c * O.pauli_expectation(n, O.run(n, O.hea_ops(n, L, tp)), s) for s, c in zip(strings, coeffs))
        em = sum(c * O.pauli_expectation(n, O.run(n, O.hea_ops(n, L, tm)), s) for s, c in zip(strings, coeffs))
        assert abs(g[k] - (ep - em) / 2e-5) < 1e-6


def _mixed_circuit(n, layers, seed):
    """Random single-qubit gate mix (with RX/RY/RZ parameters) and a CNOT ring per layer."""
    rng = np.random.default_rng(seed)
    names = ["H", "S", "T", "SX", "RX", "RY", "RZ", "X", "Y", "Z", "SDG", "TDG"]
    gates, k = [], 0
    for _ in range(layers):
        for q in range(n):
            for _ in range(2):
                name = names[rng.integers(len(names))]
                if name in ("RX", "RY", "RZ"):
                    gates.append((name, q, -1, k, float(rng.uniform(0, 2 * np.pi))))
                    k += 1
                else:
                    gates.append((name, q, -1, -1, 0.0))
        gates += [("CNOT", q, (q + 1) % n, -1, 0.0) for q in range(n)]
    return gates


@pytest.mark.parametrize("n", [5, 13, 14])
@pytest.mark.parametrize("form", [0, 1])
def test_gate_mix_both_sweep_forms(n, form):
    """Rotation form (real rotations + phase tables, pre-phase layers for general gates)
    and the general-matrix form against the oracle on a gate mix."""
    gates = _mixed_circuit(n, 3, 40 + n)
    ops, oops = Q.make_ops(gates), O.make_ops(gates)
    ref = O.qgt(n, oops)
    p = Q.Plan(n, ops, Q.MODE_MIXED, sweep_form=form)
    got, _ = p.execute()
    p.close()
    assert Q.rel_err(got, ref) <= MIXED_TOL


@pytest.mark.parametrize("n,tail", [(14, False), (15, True), (16, False)])
def test_split_plane_final_states(n, tail):
    """Depth-first plans write their final derivative states as the Gram's fp16 pieces (the
    last pass and the last block's injection), which the complex-form Gram reads: Q agrees
    with the plan that keeps complex64 states and splits them in the real-form Gram (the same
    pieces and products, summed in another order) and both match the oracle; the TFIM energy
    and gradient read back from the pieces match the oracle to 1e-6.  tail: the circuit ends
    with a rotation layer without a CNOT ring, so the final pass would run in place and is
    made to write out of place."""
    gates = _mixed_circuit(n, 2, 900 + n)
    if tail:
        k = sum(1 for g in gates if g[3] >= 0)
        rng = np.random.default_rng(n)
        gates += [("RY", q, -1, k + q, float(rng.uniform(0, 2 * np.pi))) for q in range(n)]
    ops, oops = Q.make_ops(gates), O.make_ops(gates)
    J, h = 1.0, 0.7
    out = {}
    for pre in (True, False):
        p = Q.Plan(n, ops, Q.MODE_MIXED, presplit=pre)
        p.set_hamiltonian(Q.tfim_terms(n, J, h))
        q, st = p.execute()
        assert st["presplit"] == int(pre)
        out[pre] = (q, *p.energy_gradient())
        p.close()
    ref = O.qgt(n, oops)
    err_pre, err_c64 = Q.rel_err(out[True][0], ref), Q.rel_err(out[False][0], ref)
    print(f"n={n} tail={tail}: split-plane states + complex Gram {err_pre:.2e}, complex64 + real Gram {err_c64:.2e}")
    assert err_pre <= MIXED_TOL and err_c64 <= MIXED_TOL
    assert Q.rel_err(out[True][0], out[False][0]) <= MIXED_TOL
    e_ref, g_ref = O.tfim_gradient(n, oops, J, h)
    for pre in (True, False):
        assert abs(out[pre][1] - e_ref) <= 1e-6 * max(1.0, abs(e_ref))
        assert np.max(np.abs(out[pre][2] - g_ref)) <= 1e-6 * max(1.0, np.max(np.abs(g_ref)))


@pytest.mark.parametrize("n,L", [(14, 3), (15, 4), (16, 5)])
def test_two_sided_schedule(n, L):
    """Two-sided schedule: the derivative states of late blocks are injected at their block's
    input and carried backwards through the inverse blocks (the CNOT ring undone by a gathering
    load, then u^dagger), and meet the forward-carried ones after a middle block.  Q matches
    the oracle and the all-forward schedule with fewer state-passes, and the TFIM energy and
    gradient (H psi carried back to the meeting point) match the oracle."""
    th, ops, oops = hea(n, L, seed=1300 + n)
    J, h = 1.0, 0.6
    out = {}
    for two in (True, False):
        p = Q.Plan(n, ops, Q.MODE_MIXED, two_sided=two)
        p.set_hamiltonian(Q.tfim_terms(n, J, h))
        q, st = p.execute()
        out[two] = (q, st, *p.energy_gradient())
        p.close()
    assert out[False][1]["meet_block"] == L - 1
    assert 0 <= out[True][1]["meet_block"] < L - 1
    assert out[True][1]["sweep_state_passes"] < out[False][1]["sweep_state_passes"]
    ref = O.qgt(n, oops)
    for two in (True, False):
        assert Q.rel_err(out[two][0], ref) <= MIXED_TOL, two
    assert Q.rel_err(out[True][0], out[False][0]) <= MIXED_TOL
    e_ref, g_ref = O.tfim_gradient(n, oops, J, h)
    for two in (True, False):
        assert abs(out[two][2] - e_ref) <= 1e-6 * max(1.0, abs(e_ref))
        assert np.max(np.abs(out[two][3] - g_ref)) <= 1e-6 * max(1.0, np.max(np.abs(g_ref)))


@pytest.mark.parametrize("form", [0, 1])
def test_two_sided_gate_mix_inverse_passes(form):
    """Inverse blocks of general gates (H, S, T, SX, ... factored with pre- and post-phases, or
    general complex matrices with sweep_form=1): six layers at n = 14 meet after block 2, so the
    states of blocks 4 and 5 run through one and two inverse blocks.  Q, energy and gradient
    against the oracle."""
    n = 14
    gates = _mixed_circuit(n, 6, 321 + form)
    ops, oops = Q.make_ops(gates), O.make_ops(gates)
    p = Q.Plan(n, ops, Q.MODE_MIXED, sweep_form=form)
    p.set_hamiltonian(Q.tfim_terms(n, 1.0, 0.45))
    q, st = p.execute()
    e, g = p.energy_gradient()
    p.close()
    assert 0 < st["meet_block"] < 4
    assert Q.rel_err(q, O.qgt(n, oops)) <= MIXED_TOL
    e_ref, g_ref = O.tfim_gradient(n, oops, 1.0, 0.45)
    assert abs(e - e_ref) <= 1e-6 * max(1.0, abs(e_ref))
    assert np.max(np.abs(g - g_ref)) <= 1e-6 * max(1.0, np.max(np.abs(g_ref)))


@pytest.mark.parametrize("n,L,mix", [(14, 5, False), (16, 6, False), (15, 6, True)])
def test_fused_injection_bit_identical(n, L, mix):
    """Carried derivative-state chunks are formed by their first pass while it loads the primal
    state (forward: the next block's high pass; backward: the gathering load of the inverse
    block), with the same arithmetic as the injection kernel, whose sums-only variant still
    provides <d|psi>, |d|^2, |psi|^2: Q, E and grad are bit-identical to injecting into memory
    first, and Q matches the oracle."""
    if mix:
        gates = _mixed_circuit(n, L, 555)
        ops, oops = Q.make_ops(gates), O.make_ops(gates)
    else:
        _, ops, oops = hea(n, L, seed=1700 + n)
    out = {}
    for fused in (True, False):
        p = Q.Plan(n, ops, Q.MODE_MIXED, fused_inject=fused)
        p.set_hamiltonian(Q.tfim_terms(n, 1.0, 0.8))
        q, st = p.execute()
        out[fused] = (q, *p.energy_gradient())
        p.close()
    assert np.array_equal(out[True][0], out[False][0])
    assert out[True][1] == out[False][1]
    assert np.array_equal(out[True][2], out[False][2])
    assert Q.rel_err(out[True][0], O.qgt(n, oops)) <= MIXED_TOL


def test_two_sided_meets_after_a_block_without_inverse():
    """A block whose M-stage is neither the CNOT ring nor empty (here one CNOT and one SWAP,
    applied by per-gate kernels) has no inverse passes: the meeting point moves to or past it,
    and Q, energy and gradient still match the oracle (general gates: pre-phase layers in the
    inverse passes too)."""
    n = 14
    gates = _mixed_circuit(n, 5, 77)
    # replace the ring of layer 3 (gates of a layer: 2 n single-qubit + n CNOT) by two gates
    per = 3 * n
    ring3 = slice(3 * per + 2 * n, 4 * per)
    gates[ring3] = [("CNOT", 0, 5, -1, 0.0), ("SWAP", 2, 13, -1, 0.0)]
    ops, oops = Q.make_ops(gates), O.make_ops(gates)
    p = Q.Plan(n, ops, Q.MODE_MIXED)
    p.set_hamiltonian(Q.tfim_terms(n, 1.0, 0.9))
    q, st = p.execute()
    e, g = p.energy_gradient()
    p.close()
    assert st["meet_block"] >= 3
    assert Q.rel_err(q, O.qgt(n, oops)) <= MIXED_TOL
    e_ref, g_ref = O.tfim_gradient(n, oops, 1.0, 0.9)
    assert abs(e - e_ref) <= 1e-6 * max(1.0, abs(e_ref))
    assert np.max(np.abs(g - g_ref)) <= 1e-6 * max(1.0, np.max(np.abs(g_ref)))


@pytest.mark.parametrize("n", [14, 16])
def test_lds_pass_kernel_on_register_plans(n):
    """Batched sweeps with the LDS gate-group kernel forced for full 13-bit tiles
    (the register kernel's phase-ordered gate lists), against the oracle."""
    L = 2
    th = O.uniform_angles(700 + n, 2 * n * L)
    b = Q.Batch([(n, Q.hea_ops(n, L, th))], Q.BatchTune(pass_variant=1))
    got = b.execute()[0][0]
    b.close()
    assert Q.rel_err(got, O.qgt(n, O.hea_ops(n, L, th))) <= MIXED_TOL


@pytest.mark.parametrize("method", [Q.SOLVE_CHOLESKY, Q.SOLVE_JACOBI_PCG, Q.SOLVE_CG])
def test_ng_step_device_resident(method):
    """One natural-gradient step on the device against the FP64 oracle's step: the oracle's
    Q, energy and gradient, A = Re Q + lambda I and its Cholesky solve on the host
    (update_natural_gradient, classical_optimization_engine.c:390-416).  The GPU's Q
    carries the mixed-precision error (<= 1e-6 normwise), which the solve amplifies by at
    most cond(A) <= (1/4 + lambda) / lambda: the step is held to 1e-6 cond(A) relative."""
    n, L, J, h, lam, eta = 6, 3, 1.0, 0.9, 1e-3, 0.05
    th, ops, oops = hea(n, L, seed=321)
    e_ref, g_ref = O.tfim_gradient(n, oops, J, h)
    A = O.qgt(n, oops).real + lam * np.eye(len(th))
    x_ref = O.cholesky_solve(A, g_ref)
    p = Q.Plan(n, ops, Q.MODE_MIXED)
    p.set_hamiltonian(Q.tfim_terms(n, J, h))
    theta = th.copy()
    info = Q.ng_step(p, theta, lam, eta, method, 1e-11, 20000, None, h)
    p.close()
    assert info["solve_status"] == 0
    assert abs(info["energy"] - e_ref) <= 1e-6
    assert abs(info["grad_norm"] - np.linalg.norm(g_ref)) <= 1e-6 * np.linalg.norm(g_ref)
    x = (th - theta) / eta
    cond = np.linalg.cond(A)
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(f"method {method}: step rel err {err:.2e} (cond(A) = {cond:.0f})")
    assert err <= 1e-6 * cond


@pytest.mark.parametrize("n,L", [(16, 4), (20, 6)])
def test_two_stream_chunks_match_one_stream(n, L):
    """The chain's stream carries the derivative chunks injected beyond the meeting block
    backward while the library stream carries the forward ones and the meeting point's
    injections, joined by events before the Gram: eager launches, the captured graph's replays
    and per-launch timing (every kernel on one stream, one after another) give bit-identical Q."""
    th, ops, oops = hea(n, L, seed=4500 + n)
    one = Q.Plan(n, ops, Q.MODE_MIXED, graph=False)
    Q.set_profiling(True)
    try:
        q_one, _ = one.execute()
    finally:
        Q.set_profiling(False)
    one.close()
    two = Q.Plan(n, ops, Q.MODE_MIXED)
    for want in (0, 1, 1):                       # eager, then captured and replayed
        q, st = two.execute()
        assert st["graph"] == want
        assert np.array_equal(q, q_one)
    two.close()
    if n <= 16:
        assert Q.rel_err(q_one, O.qgt(n, oops)) <= MIXED_TOL


def test_plan_graph_replays_match_eager():
    """From the second execute of an unchanged plan on, the launch sequence (two streams,
    events) runs as one captured HIP graph: every replay, with eager launches of another plan
    in between, gives Q bit-identical to the eager plan; new angles run eagerly once, are
    captured again (the executable updated in place) and match the oracle."""
    n, L = 16, 4
    th, ops, oops = hea(n, L, seed=4400)
    g = Q.Plan(n, ops, Q.MODE_MIXED)
    e = Q.Plan(n, ops, Q.MODE_MIXED, graph=False)
    q_e, st_e = e.execute()
    assert st_e["graph"] == 0
    flags = []
    for _ in range(4):
        q, st = g.execute()
        flags.append(st["graph"])
        assert np.array_equal(q, q_e)
        q2, _ = e.execute()                      # eager work between the replays
        assert np.array_equal(q2, q_e)
    assert flags == [0, 1, 1, 1]
    th2 = O.uniform_angles(4401, 2 * n * L)
    ang = Q.hea_ops(n, L, th2)["angle"]
    g.set_angles(ang)
    e.set_angles(ang)
    q_e2, _ = e.execute()
    assert Q.rel_err(q_e2, O.qgt(n, O.hea_ops(n, L, th2))) <= MIXED_TOL
    for want in (0, 1, 1):
        q, st = g.execute()
        assert st["graph"] == want
        assert np.array_equal(q, q_e2)
    g.close()
    e.close()
