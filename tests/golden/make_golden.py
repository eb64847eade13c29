"""Writes the golden fixtures of the QGT path from the FP64 C oracle.

QGTL's own tests hold no QGT value, so these vectors (pinned by the closed
forms, identities and brute force of tests/test_oracle.py) are the committed
reference for the GPU parity tests:
  cfg1_hea_n10_L4.npz   HEA n=10, L=4, P=80, theta ~ U[0,2pi) seed 0:
                        full Q, metric = Re Q, Berry = Im Q
  cfg2_block_n20_L6.npz HEA n=20, L=6, P=240, seed 0: Q on 16 rows x 16 cols
                        drawn without replacement from a seed-1 stream
Run: python tests/golden/make_golden.py   (after make -C oracle)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


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


def main():
    n, L = 10, 4
    th = O.uniform_angles(0, 2 * n * L)
    Q = O.qgt(n, O.hea_ops(n, L, th))
    np.savez(os.path.join(HERE, "cfg1_hea_n10_L4.npz"), n=n, layers=L, theta=th, Q=Q, metric=Q.real, berry=Q.imag)
    n, L = 20, 6
    P = 2 * n * L
    th = O.uniform_angles(0, P)
    idx = draw_without_replacement(1, 32, P)
    rows, cols = np.sort(idx[:16]), np.sort(idx[16:])
    B = O.qgt_block(n, O.hea_ops(n, L, th), rows, cols)
    np.savez(os.path.join(HERE, "cfg2_block_n20_L6.npz"), n=n, layers=L, theta=th, rows=rows, cols=cols, Q=B)


if __name__ == "__main__":
    main()
