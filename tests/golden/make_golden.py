"""Writes the golden fixtures of the QGT path from the FP64 C oracle.

QGTL's own tests hold no QGT value, so these vectors (pinned by the closed
forms, identities and brute force of tests/test_oracle.py) are the committed
reference for the GPU parity tests:
  cfg1_hea_n10_L4.npz   HEA n=10, L=4, P=80, theta ~ U[0,2pi) seed 0:
                        full Q, metric = Re Q, Berry = Im Q
  cfg2_block_n20_L6.npz HEA n=20, L=6, P=240, seed 0: Q on 16 rows x 16 cols
                        drawn without replacement from a seed-1 stream
  cfg4_block_n26_L8.npz HEA n=26, L=8, P=416, seed 0: Q on 16 rows x 16 cols,
                        a seed-1 splitmix64 Fisher-Yates shuffle of 0..P-1,
                        rows = sorted first 16, cols = sorted next 16
                        (about 34 GB of host memory and minutes of CPU time)
  cfg2_overlaps_n20_L6.npz  the same instance: a_k = <d_k psi|psi> for all 240
                        parameters (about 4 GB of host memory)
  ngvqe_tfim_n8_L3.npz  five natural-gradient VQE steps on the TFIM at n=8,
                        L=3 (P=48), J=1, h=0.9: per step the oracle's Q, E,
                        grad E and the Cholesky step (Re Q + lam I) x = grad E
Run: python tests/golden/make_golden.py [cfg1 cfg2 cfg2_overlaps cfg4 ngvqe]   (after make -C oracle)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle_lib as O  # noqa: E402
from quantum_geometric_tensor_amd.workloads import SplitMix64  # noqa: E402


def draw_without_replacement(seed, count, P):
    u = O.uniform_angles(seed, 4 * count) / (2 * np.pi)
    out = []
    for v in u:
        k = int(v * P)
        if k not in out:
            out.append(k)
        if len(out) == count:
            break
    return np.array(out, np.int32)


def shuffled_rows_cols(seed, P, block):
    """Fisher-Yates shuffle of 0..P-1 from a splitmix64 stream; sorted first and next `block`."""
    g = SplitMix64(seed)
    perm = list(range(P))
    for i in range(P - 1, 0, -1):
        j = int(g.uniform() * (i + 1))
        perm[i], perm[j] = perm[j], perm[i]
    return np.array(sorted(perm[:block]), np.int32), np.array(sorted(perm[block:2 * block]), np.int32)


def cfg1():
    n, L = 10, 4
    th = O.uniform_angles(0, 2 * n * L)
    Q = O.qgt(n, O.hea_ops(n, L, th))
    np.savez(os.path.join(HERE, "cfg1_hea_n10_L4.npz"), n=n, layers=L, theta=th, Q=Q, metric=Q.real, berry=Q.imag)


def cfg2():
    n, L = 20, 6
    P = 2 * n * L
    th = O.uniform_angles(0, P)
    idx = draw_without_replacement(1, 32, P)
    rows, cols = np.sort(idx[:16]), np.sort(idx[16:])
    B = O.qgt_block(n, O.hea_ops(n, L, th), rows, cols)
    np.savez(os.path.join(HERE, "cfg2_block_n20_L6.npz"), n=n, layers=L, theta=th, rows=rows, cols=cols, Q=B)


def cfg2_overlaps():
    """a_k = <d_k psi|psi> for every parameter of the n=20, L=6 instance (P=240), from the
    oracle's derivative states: the overlaps QGTL subtracts in
    quantum_geometric_tensor_network.c:1139-1166 (about 4 GB of host memory)."""
    n, L = 20, 6
    th = O.uniform_angles(0, 2 * n * L)
    psi, D = O.derivative_states(n, O.hea_ops(n, L, th))
    a = D.conj() @ psi
    np.savez(os.path.join(HERE, "cfg2_overlaps_n20_L6.npz"), n=n, layers=L, theta=th, a=a)


def cfg4():
    n, L = 26, 8
    P = 2 * n * L
    th = O.uniform_angles(0, P)
    rows, cols = shuffled_rows_cols(1, P, 16)
    B = O.qgt_block(n, O.hea_ops(n, L, th), rows, cols)
    np.savez(os.path.join(HERE, "cfg4_block_n26_L8.npz"), n=n, layers=L, theta=th, rows=rows, cols=cols, Q=B)


def ngvqe():
    """Five natural-gradient steps theta <- theta - eta x, (Re Q + lam I) x = grad E,
    every quantity from the FP64 oracle (quantum_geometric_tensor_network.c:1120-1175 for Q,
    quantum_classical_algorithms.c:320-427 for E, classical_optimization_engine.c:390-416
    for the update)."""
    n, L, J, h, lam, eta, steps = 8, 3, 1.0, 0.9, 1e-3, 0.1, 5
    P = 2 * n * L
    th = O.uniform_angles(0, P)
    theta = np.zeros((steps + 1, P))
    theta[0] = th
    Qs = np.zeros((steps, P, P), np.complex128)
    E = np.zeros(steps)
    G = np.zeros((steps, P))
    X = np.zeros((steps, P))
    for s in range(steps):
        ops = O.hea_ops(n, L, theta[s])
        Qs[s] = O.qgt(n, ops)
        E[s], G[s] = O.tfim_gradient(n, ops, J, h)
        A = Qs[s].real + lam * np.eye(P)
        X[s] = O.cholesky_solve(A, G[s])
        theta[s + 1] = theta[s] - eta * X[s]
    np.savez(os.path.join(HERE, "ngvqe_tfim_n8_L3.npz"), n=n, layers=L, J=J, h=h, lam=lam, eta=eta,
             theta=theta, Q=Qs, E=E, grad=G, step=X)


if __name__ == "__main__":
    targets = sys.argv[1:] or ["cfg1", "cfg2", "ngvqe"]
    for t in targets:
        globals()[t]()
        print(f"wrote {t}", flush=True)
