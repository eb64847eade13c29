"""The five workloads of the headline benchmark's configuration list, each run
at its full size on the GPU and checked against the FP64 oracle or the oracle's
committed fixtures (tests/golden/, written by make_golden.py).

  1  HEA n=10, L=4 (P=80): full Q, metric and Berry parts vs the fixture; the
     plain FP64 GPU path too
  2  HEA n=20, L=6 (P=240), the headline instance: a 16 x 16 oracle block, the
     overlaps a_k of all 240 parameters, and over the whole matrix Hermiticity,
     0 <= Q_kk <= 1/4 and the layer-0 closed form
  3  the batched workload, 512 instances (n in 12..18, L in 2..8) in one
     captured graph: invariants of every instance and the oracle's full Q of
     four of them, plus an oracle block of an n=18 instance
  4  HEA n=26, L=8 (P=416, 512 MB states, ~225 GB of device memory): a 16 x 16
     oracle block from the fixture, and the invariants
  5  the learned preconditioner at full size (TFIM n=16, L=6, P=192, 3 x 256
     MLP, 20k training pairs): every held-out solve converges to relative
     residual 1e-8 and agrees with the exact Cholesky solve

Error measure (the headline's): max |Q_gpu - Q_ref| / max |Q|, with max |Q|
over the whole GPU matrix (0.25, the diagonal).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
import quantum_geometric_tensor_amd as Q
from quantum_geometric_tensor_amd.workloads import hea_batch

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-6


def _layer0(n, theta):
    """Parameters before the first CNOT ring act on a product state: per qubit
    Q = [[1/4, i sin(a)/4], [-i sin(a)/4, sin(a)^2/4]] with a the RY angle."""
    a = theta[:n]
    E = np.zeros((2 * n, 2 * n), np.complex128)
    for q in range(n):
        E[q, q] = 0.25
        E[q, n + q] = 0.25j * np.sin(a[q])
        E[n + q, q] = -0.25j * np.sin(a[q])
        E[n + q, n + q] = np.sin(a[q]) ** 2 / 4
    return E


def _invariants(q, n, theta, tol=TOL):
    assert np.max(np.abs(q - q.conj().T)) <= 1e-12
    d = np.real(np.diag(q))
    assert d.min() >= -tol and d.max() <= 0.25 + tol
    assert np.max(np.abs(q[:2 * n, :2 * n] - _layer0(n, theta))) <= tol * 0.25


def test_config1_n10_L4_full_matrix():
    g = np.load(os.path.join(GOLD, "cfg1_hea_n10_L4.npz"))
    n, L, th = int(g["n"]), int(g["layers"]), g["theta"]
    assert np.array_equal(th, O.uniform_angles(0, 2 * n * L))
    ops = Q.hea_ops(n, L, th)
    Qm = Q.qgt(n, ops, Q.MODE_MIXED)
    err = Q.rel_err(Qm, g["Q"])
    print(f"config 1 mixed: rel err {err:.2e}")
    assert err <= TOL
    assert np.max(np.abs(Qm.real - g["metric"])) <= TOL * 0.25
    assert np.max(np.abs(Qm.imag - g["berry"])) <= TOL * 0.25
    Qp = Q.qgt(n, ops, Q.MODE_FP64_PLAIN)
    errp = Q.rel_err(Qp, g["Q"])
    print(f"config 1 plain FP64: rel err {errp:.2e}")
    assert errp <= 1e-11
    _invariants(Qm, n, th)


def test_config2_n20_L6_headline():
    g = np.load(os.path.join(GOLD, "cfg2_block_n20_L6.npz"))
    n, L, th = int(g["n"]), int(g["layers"]), g["theta"]
    P = 2 * n * L
    assert np.array_equal(th, O.uniform_angles(0, P))
    p = Q.Plan(n, Q.hea_ops(n, L, th), Q.MODE_MIXED)
    Qm, st = p.execute()
    a = p.overlaps()
    p.close()
    assert a.shape == (P,)
    scale = np.max(np.abs(Qm))
    rows, cols = g["rows"], g["cols"]
    err = np.max(np.abs(Qm[np.ix_(rows, cols)] - g["Q"])) / scale
    print(f"config 2 (n=20, L=6, P=240): {st['wall_ms']:.2f} ms, 16x16 oracle block rel err {err:.2e}")
    assert err <= TOL
    _invariants(Qm, n, th)
    # the overlaps that Q = G - a a^H subtracts, from the fp16x3 Gram's row of the primal state
    ga = np.load(os.path.join(GOLD, "cfg2_overlaps_n20_L6.npz"))["a"]
    aerr = np.max(np.abs(a - ga))
    print(f"config 2 overlaps: max |a_gpu - a_oracle| = {aerr:.2e} (|a| <= {np.max(np.abs(ga)):.3f})")
    assert aerr <= 1e-7
    assert np.max(np.abs(a.real)) <= 1e-7                  # a_k = (i/2) <G_k> is imaginary
    # the plain FP64 GPU path on the same block
    Qp = Q.qgt(n, Q.hea_ops(n, L, th), Q.MODE_FP64_PLAIN)
    errp = np.max(np.abs(Qp[np.ix_(rows, cols)] - g["Q"])) / scale
    print(f"config 2 plain FP64 GPU: block rel err {errp:.2e}; mixed vs plain {Q.rel_err(Qm, Qp):.2e}")
    assert errp <= 1e-11
    assert Q.rel_err(Qm, Qp) <= TOL


def test_config3_batched_512_instances():
    work = hea_batch(512, seed=0)
    b = Q.Batch([(n, Q.hea_ops(n, L, th)) for n, L, th in work])
    Qs, st = b.execute()
    print(f"config 3: 512 instances, device {st['device_ms']:.1f} ms, {st['launches']} launches, graph {st['graph']}",
          flush=True)
    assert st["graph"] == 1 and st["instances"] == 512
    for (n, L, th), q in zip(work, Qs):
        assert q.shape == (2 * n * L, 2 * n * L)
        _invariants(q, n, th)
    # the oracle's full Q of the four cheapest instances, and a block of the first n=18 one
    cost = [(1 << n) * (2 * n * L) ** 2 for n, L, _ in work]
    for i in np.argsort(cost, kind="stable")[:4]:
        n, L, th = work[i]
        err = Q.rel_err(Qs[i], O.qgt(n, O.hea_ops(n, L, th)))
        print(f"  instance {i}: n={n} L={L} rel err {err:.2e}")
        assert err <= TOL
    i = next(k for k, (n, _, _) in enumerate(work) if n == 18)
    n, L, th = work[i]
    P = 2 * n * L
    rows, cols = np.arange(0, P, max(1, P // 8))[:8], np.arange(3, P, max(1, P // 8))[:8]
    ref = O.qgt_block(n, O.hea_ops(n, L, th), rows, cols)
    err = np.max(np.abs(Qs[i][np.ix_(rows, cols)] - ref)) / np.max(np.abs(Qs[i]))
    print(f"  instance {i}: n=18 L={L} 8x8 oracle block rel err {err:.2e}")
    assert err <= TOL
    b.close()


def test_config4_n26_L8_large_instance():
    g = np.load(os.path.join(GOLD, "cfg4_block_n26_L8.npz"))
    n, L, th = int(g["n"]), int(g["layers"]), g["theta"]
    assert np.array_equal(th, O.uniform_angles(0, 2 * n * L))
    p = Q.Plan(n, Q.hea_ops(n, L, th), Q.MODE_MIXED)
    Qm, st = p.execute()
    p.close()
    rows, cols = g["rows"], g["cols"]
    err = np.max(np.abs(Qm[np.ix_(rows, cols)] - g["Q"])) / np.max(np.abs(Qm))
    print(f"config 4 (n=26, L=8, P=416): {st['wall_ms']:.0f} ms, 16x16 oracle block rel err {err:.2e}")
    assert err <= TOL
    _invariants(Qm, n, th)


def test_config5_learned_preconditioner_full_size():
    from quantum_geometric_tensor_amd import learned_pc
    r = learned_pc.run(n=16, L=6, samples=20000, hidden=256, hidden_layers=3, epochs=20, ng_runs=1, ng_steps=10,
                       log=lambda s: print(s, flush=True))
    hist = r["train"]["history"]
    assert hist[-1]["test_loss"] < hist[0]["test_loss"]     # Kaporin loss fell on held-out pairs
    ho = r["held_out"]
    for name in ("cg", "jacobi_pcg", "learned_pcg"):
        assert ho[name]["systems"] == 2000
        assert ho[name]["converged"] == ho[name]["systems"]
        assert ho[name]["true_rel_residual_max"] <= 1.05e-8
        assert ho[name]["rel_diff_vs_cholesky_max"] <= 1e-5
    assert ho["cholesky"]["status_ok"] == 2000
    print(f"held-out iterations Jacobi-PCG {ho['jacobi_pcg']['iterations_mean']:.1f}, learned-PCG "
          f"{ho['learned_pcg']['iterations_mean']:.1f} (ratio {ho['iteration_ratio_jacobi_over_learned']:.3f})")
    ng = r["ng_vqe"]
    for name in ("jacobi_pcg", "learned_pcg", "cholesky"):
        assert np.isfinite(ng[name]["final_energies"]).all()
    print(f"NG steps/s: Jacobi-PCG {ng['jacobi_pcg']['steps_per_s']:.0f}, learned-PCG "
          f"{ng['learned_pcg']['steps_per_s']:.0f}, Cholesky {ng['cholesky']['steps_per_s']:.0f}")
