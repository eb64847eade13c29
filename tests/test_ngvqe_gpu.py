"""The natural-gradient VQE driver (qgtl_ng_step) against the FP64 oracle's
trajectory: five steps theta <- theta - eta x, (Re Q + lambda I) x = grad E,
on the TFIM at n=8, L=3 (P=48), J=1, h=0.9, lambda=1e-3, eta=0.1, every
quantity of the oracle trajectory computed in FP64 on the host
(tests/golden/ngvqe_tfim_n8_L3.npz, written by make_golden.py; QGTL's caller
is update_natural_gradient, classical_optimization_engine.c:390-416).

Per step from the oracle's own angles: E to 1e-6, grad E to 1e-6 relative,
the step x to 1e-6 cond(A) relative (the mixed-precision Q error, <= 1e-6
normwise, amplified by at most the condition number of A).  Free running from
theta_0: the five energies and the final angles follow the oracle's.
"""
import os

import numpy as np
import pytest

import quantum_geometric_tensor_amd as Q

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ngvqe_tfim_n8_L3.npz")


def _load():
    g = np.load(GOLD)
    return {k: g[k] for k in g.files}


@pytest.mark.parametrize("method", [Q.SOLVE_CHOLESKY, Q.SOLVE_JACOBI_PCG])
def test_ng_steps_from_oracle_angles(method):
    g = _load()
    n, L, J, h, lam, eta = int(g["n"]), int(g["layers"]), float(g["J"]), float(g["h"]), float(g["lam"]), float(g["eta"])
    plan = Q.Plan(n, Q.hea_ops(n, L, g["theta"][0]), Q.MODE_MIXED)
    plan.set_hamiltonian(Q.tfim_terms(n, J, h))
    for s in range(len(g["E"])):
        # energy and gradient at the oracle's angles
        plan.set_angles(Q.hea_ops(n, L, g["theta"][s])["angle"])
        Qm, _ = plan.execute()
        e, grad = plan.energy_gradient()
        assert abs(e - g["E"][s]) <= 1e-6
        assert np.linalg.norm(grad - g["grad"][s]) <= 1e-6 * np.linalg.norm(g["grad"][s])
        assert Q.rel_err(Qm, g["Q"][s]) <= 1e-6
        # the device-resident step
        theta = g["theta"][s].copy()
        info = Q.ng_step(plan, theta, lam, eta, method, 1e-12, 20000)
        assert info["solve_status"] == 0
        x = (g["theta"][s] - theta) / eta
        A = g["Q"][s].real + lam * np.eye(len(x))
        cond = np.linalg.cond(A)
        err = np.linalg.norm(x - g["step"][s]) / np.linalg.norm(g["step"][s])
        print(f"step {s}: E {info['energy']:.8f} (oracle {g['E'][s]:.8f}), step rel err {err:.2e}, cond {cond:.0f}")
        assert abs(info["energy"] - g["E"][s]) <= 1e-6
        assert err <= 1e-6 * cond
    plan.close()


def test_ng_trajectory_free_running():
    g = _load()
    n, L, J, h, lam, eta = int(g["n"]), int(g["layers"]), float(g["J"]), float(g["h"]), float(g["lam"]), float(g["eta"])
    plan = Q.Plan(n, Q.hea_ops(n, L, g["theta"][0]), Q.MODE_MIXED)
    plan.set_hamiltonian(Q.tfim_terms(n, J, h))
    theta = g["theta"][0].copy()
    for s in range(len(g["E"])):
        info = Q.ng_step(plan, theta, lam, eta, Q.SOLVE_CHOLESKY)
        print(f"step {s}: E {info['energy']:.8f} oracle {g['E'][s]:.8f}; "
              f"|theta - theta_oracle| {np.max(np.abs(theta - g['theta'][s + 1])):.2e}")
        assert abs(info["energy"] - g["E"][s]) <= 1e-5 * max(1.0, abs(g["E"][s]))
    assert np.max(np.abs(theta - g["theta"][-1])) <= 1e-4
    plan.close()
