"""CPU checks of the C-ABI boundary (no GPU work is launched).

- every function declared in include/*.h is exported by libqgtl_mi355x.so;
- the structs and enum values of the drop-in API have QGTL's layout
  (sizeof/offsetof printed by oracle/abi_layout.c against QGTL's headers and
  against include/; skipped when the QGTL build of the probe is absent);
- host-side plan validation returns QGTL's error codes;
- the host HEA builder matches the oracle's.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import quantum_geometric_tensor_amd as Q

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)     # preprocessor lines (macros with arguments)
    text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)   # struct/enum typedefs
    names = set()
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_\s\*]*?)\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        ret, name = m.group(1).strip(), m.group(2)
        if not ret or name in ("if", "while", "for", "sizeof") or ret.startswith("#"):
            continue
        names.add(name)
    return names


@pytest.mark.parametrize("header", ["qgtl_api.h", "qgtl_mi355x.h", "qgtl_hip_abi.h", "qgtl_compute_backend.h"])
def test_all_declared_symbols_exported(header):
    lib = Q.lib()
    names = declared_functions(os.path.join(INCLUDE, header))
    assert len(names) >= 2
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_abi_layout_matches_qgtl_headers():
    ref = os.path.join(ROOT, "oracle", "_ref", "abi_layout_ref")
    ours = os.path.join(ROOT, "tests", "c", "_build", "abi_layout_ours")
    if not os.path.exists(ref):
        pytest.skip("QGTL-header build of the ABI probe not present (needs the QGTL sources)")
    a = subprocess.run([ref], stdout=subprocess.PIPE, text=True, check=True).stdout
    b = subprocess.run([ours], stdout=subprocess.PIPE, text=True, check=True).stdout
    assert a.count("\n") > 80
    assert a == b


def test_compute_backend_registers_with_qgtl_registry():
    """Loading the library registers the HIP backend through QGTL's compute_register_backend
    (here a stand-in defined by the test program), once, with every vtable slot filled."""
    prog = os.path.join(ROOT, "tests", "c", "_build", "test_backend_register")
    res = subprocess.run([prog], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert res.returncode == 0, res.stdout
    assert "HIP" in res.stdout


def test_register_pass_descriptors_match_dense_blocks():
    """Every register pass the plan builder emits, applied in FP64 on the host by the
    descriptor contract of include/qgtl_hip_abi.h (single-form K(t) gates, the store at
    j ^ xmask, the fold signs, the phase layers moved between a block's two passes), composes
    to each block and each inverse block applied densely (tests/c/test_pass_form.c): HEA and
    RX-RY-RZ ansatze at 14, 15 and 17 qubits, max |diff| < 2e-5 (FP32 descriptor values)."""
    prog = os.path.join(ROOT, "tests", "c", "_build", "test_pass_form")
    res = subprocess.run([prog], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert res.returncode == 0, res.stdout
    assert res.stdout.strip().endswith("ok")


def test_plan_validation_errors():
    lib = Q.lib()
    bad_qubit = Q.make_ops([("RY", 5, -1, 0, 0.1)])
    h = lib.qgtl_plan_create(3, 1, bad_qubit.ctypes.data_as(ctypes.c_void_p), 0, None)
    assert not h and b"out of range" in lib.qgtl_last_error()
    bad_param = Q.make_ops([("H", 0, -1, 0, 0.0)])
    h = lib.qgtl_plan_create(2, 1, bad_param.ctypes.data_as(ctypes.c_void_p), 0, None)
    assert not h and b"RX/RY/RZ" in lib.qgtl_last_error()
    dup = Q.make_ops([("RY", 0, -1, 0, 0.1), ("RZ", 1, -1, 0, 0.2)])
    h = lib.qgtl_plan_create(2, 2, dup.ctypes.data_as(ctypes.c_void_p), 0, None)
    assert not h and b"each once" in lib.qgtl_last_error()
    cz = Q.make_ops([("RY", 0, -1, 0, 0.1), ("CZ", 0, 0, -1, 0.0)])
    h = lib.qgtl_plan_create(2, 2, cz.ctypes.data_as(ctypes.c_void_p), 0, None)
    assert not h
    unsupported = np.zeros(1, Q.OP_DTYPE)
    unsupported[0] = (33, 0, 1, -1, 0.0)          # GATE_TYPE_XX
    h = lib.qgtl_plan_create(2, 1, unsupported.ctypes.data_as(ctypes.c_void_p), 0, None)
    assert not h and b"not supported" in lib.qgtl_last_error()
    good = Q.hea_ops(4, 2, np.linspace(0, 1, 16))
    p = Q.Plan(4, good)           # plan creation allocates nothing on the device
    assert p.P == 16
    p.close()


def test_hea_builder_matches_oracle():
    for n, L in [(1, 2), (2, 3), (7, 2)]:
        th = O.uniform_angles(n, 2 * n * L)
        a = Q.hea_ops(n, L, th)
        b = O.hea_ops(n, L, th)
        assert a.tobytes() == b.tobytes()


def test_batch_validation_errors():
    """Batch creation checks every instance on the host before touching the GPU."""
    ring_free = Q.make_ops([("RY", 0, -1, 0, 0.3), ("CNOT", 0, 1, -1, 0.0), ("RY", 1, -1, 1, 0.2)])
    with pytest.raises(Q.QGTLError, match="ring"):
        Q.Batch([(3, Q.hea_ops(3, 1, np.ones(6))), (14, ring_free)])
    no_params = Q.make_ops([("H", 0, -1, -1, 0.0)])
    with pytest.raises(Q.QGTLError, match="no parameters"):
        Q.Batch([(2, no_params)])
    with pytest.raises(Q.QGTLError, match="instance 1"):
        Q.Batch([(2, Q.hea_ops(2, 1, np.ones(4))), (2, Q.make_ops([("RY", 4, -1, 0, 0.1)]))])


def _form(plan):
    err = ctypes.c_double(0.0)
    npre = ctypes.c_int(0)
    lib = Q.lib()
    lib.qgtl_plan_debug_form.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    assert lib.qgtl_plan_debug_form(plan._h, ctypes.byref(err), ctypes.byref(npre)) == 0
    return err.value, npre.value


def test_rotation_form_reconstructs_every_gate():
    """Sweep passes apply each qubit's 2x2 product as diag(post) [[c,-s],[s,c]] diag(pre);
    the factors reproduce the general matrices to fp32 rounding (host only)."""
    rng = np.random.default_rng(9)
    # hardware-efficient ansatz: RZ.RY products need no pre-phase layer
    for n in (3, 13, 16):
        p = Q.Plan(n, Q.hea_ops(n, 3, rng.uniform(0, 2 * np.pi, 6 * n)))
        err, npre = _form(p)
        assert err <= 2e-7 and npre == 0, (n, err, npre)
        p.close()
    # general gate mix: H, S, T, SX, RX, Y, Z, PHASE products, anti-diagonal and diagonal cases
    names = ["H", "S", "T", "SX", "RX", "RY", "RZ", "X", "Y", "Z", "SDG", "TDG"]
    for n in (4, 14):
        gates, k = [], 0
        for layer in range(4):
            for q in range(n):
                for _ in range(3):
                    name = names[rng.integers(len(names))]
                    if name in ("RX", "RY", "RZ"):
                        gates.append((name, q, -1, k, float(rng.uniform(0, 2 * np.pi))))
                        k += 1
                    else:
                        gates.append((name, q, -1, -1, 0.0))
            gates += [("CNOT", q, (q + 1) % n, -1, 0.0) for q in range(n)]
        p = Q.Plan(n, Q.make_ops(gates))
        err, npre = _form(p)
        assert err <= 4e-7, (n, err)
        p.close()
    # lone X and Z: anti-diagonal and diagonal matrices
    p = Q.Plan(2, Q.make_ops([("X", 0, -1, -1, 0.0), ("Z", 1, -1, -1, 0.0), ("RY", 0, -1, 0, 0.3)]))
    assert _form(p)[0] <= 2e-7
    p.close()
