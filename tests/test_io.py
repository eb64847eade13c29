"""Files of the QGT path (no GPU needed): computed QGTs and learned-preconditioner
training sets round-trip bit for bit, and the loader rejects what it must:
another kind of file, a flipped payload byte (checksum), a truncated file, a
size the caller did not expect."""
import os

import numpy as np
import pytest

import quantum_geometric_tensor_amd as Q


def _hermitian(P, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(P, P)) + 1j * rng.normal(size=(P, P))
    return A @ A.conj().T


def test_qgt_round_trip(tmp_path):
    P = 37
    Qm = _hermitian(P, 1)
    th = np.random.default_rng(2).uniform(0, 2 * np.pi, P)
    f = tmp_path / "q.qgt"
    Q.qgt_save(f, 7, Qm, th)
    n, Q2, th2 = Q.qgt_load(f)
    assert n == 7
    assert np.array_equal(Q2, Qm) and np.array_equal(th2, th)
    Q.qgt_save(f, 7, Qm)                              # without angles
    n, Q3, th3 = Q.qgt_load(f)
    assert th3 is None and np.array_equal(Q3, Qm)


def test_dataset_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    X = rng.normal(size=(11, 25)).astype(np.float32)
    As = rng.normal(size=(11, 8, 8)).astype(np.float32)
    f = tmp_path / "s.set"
    Q.dataset_save(f, X, As)
    X2, As2 = Q.dataset_load(f)
    assert np.array_equal(X2, X) and np.array_equal(As2, As)


def test_rejections(tmp_path):
    P = 9
    f = tmp_path / "q.qgt"
    Q.qgt_save(f, 3, _hermitian(P, 4))
    raw = bytearray(f.read_bytes())
    # another kind of file
    g = tmp_path / "s.set"
    Q.dataset_save(g, np.zeros((2, 3), np.float32), np.zeros((2, 4, 4), np.float32))
    with pytest.raises(Q.QGTLError) as e:
        Q.qgt_load(g)
    assert e.value.code == -5
    # one payload byte flipped -> checksum
    bad = bytearray(raw)
    bad[-3] ^= 0x10
    h = tmp_path / "bad.qgt"
    h.write_bytes(bytes(bad))
    with pytest.raises(Q.QGTLError) as e:
        Q.qgt_load(h)
    assert e.value.code == -5 and "checksum" in str(e.value)
    # truncated
    h.write_bytes(bytes(raw[:-8]))
    with pytest.raises(Q.QGTLError) as e:
        Q.qgt_load(h)
    assert e.value.code in (-5, -16)
    # trailing garbage
    h.write_bytes(bytes(raw) + b"x")
    with pytest.raises(Q.QGTLError) as e:
        Q.qgt_load(h)
    assert e.value.code == -5
    # size the caller does not expect
    L = Q._io_lib()
    buf = np.zeros((P + 1, P + 1), np.complex128)
    assert L.qgtl_qgt_load(os.fsencode(f), P + 1, None, Q._p(buf)) == -3
    # missing file
    with pytest.raises(Q.QGTLError) as e:
        Q.qgt_load(tmp_path / "nope.qgt")
    assert e.value.code == -16
