"""GPU tests of the rest of the drop-in boundary, through tests/c/test_dropin_more
(QGTL's C API called from C, as a QGTL program calls it):

- flat circuits built the way QGTL's circuit builders build them
  (quantum_circuit_operations.c:786-1135: only type, num_qubits, qubits[] and
  parameters set) and applied with apply_quantum_circuit: Q to 1e-6 normwise
  and the state to 1e-5 (up to a global phase) of the FP64 oracle;
- two host threads computing full QGTs of two same-structure networks at
  different angles at the same time: each result equals its oracle QGT and its
  own earlier repetitions bit for bit;
- the quantum_state_* set on the host and on the device against numpy;
- geometric_tensor_save / geometric_tensor_load round trip and rejections;
- the HIP ComputeBackendOps slots against numpy in FP64.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "c", "_build", "test_dropin_more")


def _run(*args, timeout=120):
    res = subprocess.run([PROG, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=timeout)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stdout


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_flat_circuit_from_qgtl_builders(tmp_path):
    out = tmp_path / "flat.bin"
    _run("flat", out)
    d = np.fromfile(out, np.float64)
    P = int(d[0])
    Qg = d[1:1 + 2 * P * P].view(np.complex128).reshape(P, P)
    psi = d[1 + 2 * P * P:].view(np.complex128)
    a = [0.3, 1.1, -0.7, 2.9, 0.45, -1.8, 0.05, 3.7]
    gates = [("H", 0, -1, -1, 0.0), ("RY", 1, -1, 0, a[0]), ("CNOT", 0, 1, -1, 0.0), ("RX", 2, -1, 1, a[1]),
             ("X", 3, -1, -1, 0.0), ("RZ", 3, -1, 2, a[2]), ("CZ", 2, 4, -1, 0.0), ("RY", 4, -1, 3, a[3]),
             ("S", 0, -1, -1, 0.0), ("SWAP", 1, 3, -1, 0.0), ("Y", 2, -1, -1, 0.0), ("RX", 0, -1, 4, a[5]),
             ("CNOT", 4, 0, -1, 0.0), ("Z", 1, -1, -1, 0.0), ("RZ", 4, -1, 5, a[6]), ("RY", 2, -1, 6, a[7])]
    ops = O.make_ops(gates)
    Qref = O.qgt(5, ops)
    assert P == 7
    err = _rel(Qg, Qref)
    print(f"flat circuit: P={P} rel err {err:.2e}")
    assert err <= 1e-6
    ref = O.run(5, ops)
    ov = np.vdot(ref, psi)
    assert abs(abs(ov) - 1.0) < 1e-5                      # the same state up to a global phase
    assert np.max(np.abs(psi * np.conj(ov) / abs(ov) - ref)) < 1e-5


def test_concurrent_same_structure_networks(tmp_path):
    n, L, sa, sb = 10, 4, 11, 12
    out = tmp_path / "thr.bin"
    _run("threads", n, L, sa, sb, 4, out)
    d = np.fromfile(out, np.float64)
    P = int(d[0])
    assert P == 2 * n * L
    Q2 = d[1:].view(np.complex128).reshape(2, P, P)
    for seed, Qt in zip((sa, sb), Q2):
        ref = O.qgt(n, O.hea_ops(n, L, O.uniform_angles(seed, P)))
        err = _rel(Qt, ref)
        print(f"thread seed {seed}: rel err {err:.2e}")
        assert err <= 1e-6
    assert _rel(Q2[0], Q2[1]) > 1e-2                      # different angles, different QGTs


def test_state_container_host_and_device(tmp_path):
    out = tmp_path / "state.bin"
    _run("state", out)
    d = np.fromfile(out, np.float64)
    N = int(d[0])
    o = 1
    amp = d[o:o + 2 * N].view(np.complex128); o += 2 * N
    amp2 = d[o:o + 2 * N].view(np.complex128); o += 2 * N
    ip_h = d[o] + 1j * d[o + 1]; fid_h = d[o + 2]; norm_h = d[o + 3]; o += 4
    ip_d = d[o] + 1j * d[o + 1]; ip_dd = d[o + 2] + 1j * d[o + 3]; fid_d = d[o + 4]; norm_d = d[o + 5]; o += 6
    a_final = d[o:o + 2 * N].view(np.complex128); o += 2 * N
    ipb = d[o] + 1j * d[o + 1]
    # QGTL's initialize: FP64 norm, FP32 scale when |norm^2 - 1| >= 1e-6
    a32 = amp.astype(np.complex64)
    ns = float(np.sum(np.abs(amp) ** 2))
    a_init = (a32 * np.float32(1.0 / np.sqrt(np.float32(ns)))).astype(np.complex128)
    b_init = (amp2.astype(np.complex64) * np.float32(1.0 / np.sqrt(np.float32(np.sum(np.abs(amp2) ** 2))))
              ).astype(np.complex128)
    ip = np.vdot(a_init, b_init)
    for got in (ip_h, ip_d, ip_dd):
        assert abs(got - ip) <= 1e-6 * max(1.0, abs(ip))
    assert abs(fid_h - abs(ip) ** 2) < 1e-6 and abs(fid_d - abs(ip) ** 2) < 1e-6
    assert abs(norm_h - 1.0) < 1e-6 and abs(norm_d - 1.0) < 1e-6
    # normalize on the device of the raw amp2 values, read back with from_device
    want = amp2 / np.linalg.norm(amp2)
    assert np.max(np.abs(a_final - want)) < 1e-6
    # <a_final|e_5> = conj(a_final[5])
    assert abs(ipb - np.conj(a_final[5])) < 1e-6


def test_geometric_tensor_save_load(tmp_path):
    _run("tensor", tmp_path)


def test_compute_backend_slots(tmp_path):
    out = tmp_path / "backend.bin"
    _run("backend", out)
    d = np.fromfile(out, np.float64)
    N = int(d[0])
    m, nn, kk = 5, 7, 3
    o = 1

    def take(count):
        nonlocal o
        v = d[o:o + count]
        o += count
        return v

    psi = take(2 * N).view(np.complex128)
    phi = take(2 * N).view(np.complex128)
    U = take(2 * N * N).view(np.complex128).reshape(N, N)
    obs = take(N)
    A = take(2 * m * nn).view(np.complex128).reshape(m, nn)
    B = take(2 * nn * kk).view(np.complex128).reshape(nn, kk)
    r = take(2)
    g = take(2)
    e = take(1)[0]
    Upsi = take(2 * N).view(np.complex128)
    nrm = take(2 * N).view(np.complex128)
    C = take(2 * m * kk).view(np.complex128).reshape(m, kk)
    rd = take(2)
    ip = np.vdot(psi, phi)
    assert abs(complex(r[0], r[1]) - ip) <= 1e-6 * max(1.0, abs(ip))
    gr = np.vdot(phi, psi)                                 # <backward|forward>
    assert abs(complex(g[0], g[1]) - gr) <= 1e-6 * max(1.0, abs(gr))
    ex = float(np.sum(np.abs(psi) ** 2 * obs))
    assert abs(e - ex) <= 1e-6 * max(1.0, abs(ex))
    assert _rel(Upsi, U @ psi) <= 1e-6                     # row-major U, FP64 accumulation
    assert _rel(nrm, phi / np.linalg.norm(phi)) <= 1e-6
    assert _rel(C, A @ B) <= 1e-5                          # FP32 accumulation, as QGTL's CPU backend
    up = U @ psi
    assert abs(complex(rd[0], rd[1]) - np.vdot(up, up)) <= 1e-5 * abs(np.vdot(up, up))
