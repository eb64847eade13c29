"""Pins of the FP64 CPU oracle of the learned preconditioner (no GPU):
the Kaporin loss against a dense evaluation, its scale invariance, and the
oracle's analytic gradients against central finite differences."""
import numpy as np

import oracle_lib as O


def _spd_unit_diag(P, seed):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(P, P + 3))
    A = B @ B.T / P + 1e-2 * np.eye(P)
    d = 1 / np.sqrt(np.diag(A))
    return A * d[:, None] * d[None, :]


def _dense_kaporin(P, r, y, As):
    y = y.reshape(P, 1 + r)
    s, u = y[:, 0], y[:, 1:] / np.sqrt(P)
    M = np.diag(np.exp(s)) + u @ u.T
    return P * np.log(np.trace(M @ As) / P) - np.linalg.slogdet(M)[1]


def test_kaporin_matches_dense_evaluation():
    rng = np.random.default_rng(0)
    for P, r in [(5, 0), (6, 2), (12, 4)]:
        As = _spd_unit_diag(P, P)
        y = rng.normal(scale=0.5, size=P * (1 + r))
        assert abs(O.kaporin(P, r, y, As, grad=False) - _dense_kaporin(P, r, y, As)) < 1e-11


def test_kaporin_scale_invariant_and_minimal_at_inverse():
    P, r = 8, 3
    As = _spd_unit_diag(P, 3)
    rng = np.random.default_rng(1)
    y = rng.normal(scale=0.4, size=(P, 1 + r))
    c = 0.7
    y2 = y.copy()
    y2[:, 0] += c
    y2[:, 1:] *= np.exp(c / 2)
    assert abs(O.kaporin(P, r, y.ravel(), As, grad=False) - O.kaporin(P, r, y2.ravel(), As, grad=False)) < 1e-11
    # log of Kaporin's number (tr/P)^P / det of Mt As is >= 0 (AM-GM) once -log det As is restored
    base = O.kaporin(P, r, y.ravel(), As, grad=False) - np.linalg.slogdet(As)[1]
    assert base >= -1e-12


def test_kaporin_gradient_finite_differences():
    P, r = 7, 3
    As = _spd_unit_diag(P, 5)
    y = np.random.default_rng(2).normal(scale=0.5, size=P * (1 + r))
    _, dy = O.kaporin(P, r, y, As)
    h = 1e-6
    for i in range(y.size):
        e = np.zeros_like(y)
        e[i] = h
        fd = (O.kaporin(P, r, y + e, As, grad=False) - O.kaporin(P, r, y - e, As, grad=False)) / (2 * h)
        assert abs(fd - dy[i]) < 1e-7 * max(1.0, abs(fd)), (i, fd, dy[i])


def test_mlp_gradient_finite_differences():
    F, H, nh, P, r, B = 5, 6, 3, 4, 2, 3
    Od = P * (1 + r)
    n = O.mlp_num_params(F, H, nh, Od)
    assert n == F * H + H + 2 * (H * H + H) + H * Od + Od
    rng = np.random.default_rng(3)
    params = rng.normal(scale=0.3, size=n)
    X = rng.normal(size=(B, F))
    As = np.stack([_spd_unit_diag(P, 10 + b) for b in range(B)])
    loss, g = O.mlp_loss_grad(F, H, nh, P, r, params, X, As)
    h = 1e-6
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = h
        fd = (O.mlp_loss_grad(F, H, nh, P, r, params + e, X, As, grad=False) -
              O.mlp_loss_grad(F, H, nh, P, r, params - e, X, As, grad=False)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # forward agrees with the loss path
    Y = O.mlp_forward(F, H, nh, Od, params, X)
    assert abs(np.mean([O.kaporin(P, r, Y[b], As[b], grad=False) for b in range(B)]) - loss) < 1e-12
