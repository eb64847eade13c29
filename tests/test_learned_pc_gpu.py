"""The learned preconditioner end to end on the GPU, at a small size:
training pairs from the batched QGT and from NG-VQE trajectories (TFIM,
n=8, L=4, P=64), training by the HIP kernels, then PCG on held-out systems
(Re Q + 1e-3 I) x = grad E to relative residual 1e-8.

Every iterative solution is checked against the exact FP64 Cholesky solve and
its true residual in FP64 on the host.  The gain is what the trained network
saves over the method without it (Jacobi-PCG) in PCG iterations on held-out
systems; at this size and seed it is 1.14x (47.0 vs 53.6 iterations).
"""
import numpy as np
import pytest

import quantum_geometric_tensor_amd as Q
from quantum_geometric_tensor_amd import learned_pc

pytestmark = pytest.mark.gpu


def test_learned_preconditioner_small_end_to_end(tmp_path):
    prefix = str(tmp_path / "pc")
    r = learned_pc.run(n=8, L=4, samples=1000, epochs=15, traj_steps=100, ng_runs=1, ng_steps=5,
                       log=lambda s: print(s, flush=True), save_prefix=prefix)
    # the saved training set and network reload (shapes of the n=8, L=4 workload: P = 64)
    X, As = Q.dataset_load(prefix + ".set")
    assert X.shape == (1000, 3 * 64 + 1) and As.shape == (1000, 64, 64)
    m = Q.MLP.load(prefix + ".mlp")
    assert (m.in_dim, m.P, m.rank) == (3 * 64 + 1, 64, 8)
    m.close()
    hist = r["train"]["history"]
    assert hist[-1]["test_loss"] < hist[0]["test_loss"] - 1.0          # Kaporin loss fell on held-out data
    ho = r["held_out"]
    for name in ("cg", "jacobi_pcg", "learned_pcg"):
        assert ho[name]["converged"] == ho[name]["systems"]
        assert ho[name]["true_rel_residual_max"] <= 1.05e-8
        assert ho[name]["rel_diff_vs_cholesky_max"] <= 1e-5
    gain = ho["iteration_ratio_jacobi_over_learned"]
    print(f"held-out PCG iterations: Jacobi {ho['jacobi_pcg']['iterations_mean']:.1f}, "
          f"learned {ho['learned_pcg']['iterations_mean']:.1f} (gain {gain:.3f}x)")
    assert gain >= 1.0
    ng = r["ng_vqe"]
    assert ng["learned_pcg"]["fallbacks"] == 0
    for name in ("jacobi_pcg", "learned_pcg", "cholesky"):
        assert np.isfinite(ng[name]["final_energies"]).all()
