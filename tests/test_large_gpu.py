"""Large single instances (depth-first schedule with states wider than the
Infinity Cache) against the FP64 CPU oracle on random sub-blocks, plus the
exact layer-0 closed form, Hermiticity and 0 <= Q_kk <= 1/4 over the whole
matrix.

The n=26, L=8 (P=416) instance (~225 GB of device memory) is checked against
a committed 16 x 16 oracle block in test_configs_gpu.py; the n=22 case here
runs the same code path against the oracle computed on the spot.
"""
import numpy as np
import pytest

import oracle_lib as O
import quantum_geometric_tensor_amd as Q
from quantum_geometric_tensor_amd.workloads import SplitMix64

pytestmark = pytest.mark.gpu


def _layer0(n, theta):
    a = theta[:n]
    E = np.zeros((2 * n, 2 * n), np.complex128)
    for q in range(n):
        E[q, q] = 0.25
        E[q, n + q] = 0.25j * np.sin(a[q])
        E[n + q, q] = -0.25j * np.sin(a[q])
        E[n + q, n + q] = np.sin(a[q]) ** 2 / 4
    return E


def _check(n, L, block, seed_rows=1):
    P = 2 * n * L
    theta = O.uniform_angles(0, P)
    ops = Q.hea_ops(n, L, theta)
    p = Q.Plan(n, ops, Q.MODE_MIXED)
    Qm, st = p.execute()
    p.close()
    print(f"n={n} L={L} P={P}: wall {st['wall_ms']:.1f} ms", flush=True)
    assert np.max(np.abs(Qm - Qm.conj().T)) <= 1e-12
    d = np.real(np.diag(Qm))
    assert d.min() >= -1e-6 and d.max() <= 0.25 + 1e-6
    assert np.max(np.abs(Qm[:2 * n, :2 * n] - _layer0(n, theta))) <= 0.25e-6
    # random sub-block, rows and columns drawn without replacement from seed 1
    g = SplitMix64(seed_rows)
    perm = list(range(P))
    for i in range(P - 1, 0, -1):
        j = int(g.uniform() * (i + 1))
        perm[i], perm[j] = perm[j], perm[i]
    rows = np.array(sorted(perm[:block]), np.int32)
    cols = np.array(sorted(perm[block:2 * block]), np.int32)
    ref = O.qgt_block(n, O.hea_ops(n, L, theta), rows, cols)
    sub = Qm[np.ix_(rows, cols)]
    err = np.max(np.abs(sub - ref)) / np.max(np.abs(Qm))
    print(f"  {block}x{block} oracle sub-block: normwise err {err:.2e}", flush=True)
    assert err <= 1e-6


def test_depth_first_n22():
    _check(22, 3, 8)
