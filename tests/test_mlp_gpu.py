"""HIP training and inference kernels of the learned preconditioner against
the FP64 CPU oracle (oracle/mlp_oracle.c).

The GPU rounds weights and every GEMM operand (activations, back-propagated
gradients) to bf16 and accumulates in fp32; the oracle gets the same
bf16-rounded weights and inputs and computes everything else in FP64.  Bars:
forward output normwise 6e-3, loss 1e-2 relative, parameter gradient
normwise 1e-2 with cosine similarity >= 0.9999, one Adam step to 1e-6 (the bars
are about 3x the measured bf16 errors: forward 2.0e-3, gradient 3.3e-3,
cosine 0.999997).
"""
import numpy as np
import pytest

import oracle_lib as O
import quantum_geometric_tensor_amd as Q

pytestmark = pytest.mark.gpu


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def spd_unit_diag(P, seed):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(P, P + 2))
    A = B @ B.T / P + 3e-2 * np.eye(P)
    d = 1 / np.sqrt(np.diag(A))
    return A * d[:, None] * d[None, :]


def oracle_params(m):
    """GPU master parameters with the weights rounded to bf16 (biases stay fp32)."""
    p = m.get_params().astype(np.float64)
    out = p.copy()
    o = 0
    for i in range(m.hidden_layers + 1):
        fin = m.in_dim if i == 0 else m.hidden
        fout = m.out_dim if i == m.hidden_layers else m.hidden
        out[o:o + fin * fout] = bf16(p[o:o + fin * fout])
        o += fin * fout + fout
    assert o == p.size
    return out


def make_net(seed=3, F=19, H=48, nh=3, P=6, r=2):
    m = Q.MLP(F, H, nh, P, r, seed=seed)
    # larger output weights than the Jacobi-like start, so the check sees a non-trivial preconditioner
    p = m.get_params()
    p *= 4.0
    m.set_params(p)
    return m


def test_mlp_forward_matches_oracle():
    m = make_net()
    X = np.random.default_rng(0).normal(size=(37, m.in_dim)).astype(np.float32)
    Y = m.forward(X).astype(np.float64)
    Yo = O.mlp_forward(m.in_dim, m.hidden, m.hidden_layers, m.out_dim, oracle_params(m), bf16(X))
    err = np.max(np.abs(Y - Yo)) / np.max(np.abs(Yo))
    print(f"forward normwise err {err:.2e}")
    assert err <= 6e-3
    m.close()


# gradient bars about 3x the measured normwise errors (3.3e-3 at rank 2, 1.1e-2 at rank 32: the network's
# activations enter each GEMM as bf16, the oracle keeps them in FP64)
@pytest.mark.parametrize("P,r,gbar", [(6, 2, 1e-2), (40, 32, 3.5e-2)])
def test_mlp_loss_and_gradient_match_oracle(P, r, gbar):
    m = make_net(seed=5, P=P, r=r)
    S = 20
    rng = np.random.default_rng(1)
    X = rng.normal(size=(S, m.in_dim)).astype(np.float32)
    As = np.stack([spd_unit_diag(m.P, 100 + s) for s in range(S)])
    m.set_data(X, As.astype(np.float32))
    idx = np.array([3, 7, 0, 19, 11, 5, 5, 12], np.int32)
    loss, g = m.loss_grad(idx)
    lo, go = O.mlp_loss_grad(m.in_dim, m.hidden, m.hidden_layers, m.P, m.rank, oracle_params(m), bf16(X[idx]),
                             bf16(As[idx]))
    gerr = np.max(np.abs(g - go)) / np.max(np.abs(go))
    cos = float(np.dot(g, go) / (np.linalg.norm(g) * np.linalg.norm(go)))
    print(f"loss {loss:.6f} oracle {lo:.6f}; grad normwise err {gerr:.2e} cos {cos:.6f}")
    assert abs(loss - lo) <= 1e-2 * max(1.0, abs(lo))
    assert gerr <= gbar and cos >= 0.9999
    m.close()


def test_adam_step_matches_formula():
    m = make_net(seed=7, F=11, H=32, P=5, r=1)
    S = 6
    rng = np.random.default_rng(2)
    m.set_data(rng.normal(size=(S, m.in_dim)).astype(np.float32),
               np.stack([spd_unit_diag(m.P, s) for s in range(S)]).astype(np.float32))
    idx = np.arange(S, dtype=np.int32)
    w0 = m.get_params().astype(np.float64)
    _, g = m.loss_grad(idx)
    lr = 1e-3
    m.train_step(idx, lr)
    w1 = m.get_params().astype(np.float64)
    g = g.astype(np.float64)
    mhat = (0.1 * g) / (1 - 0.9)
    vhat = (0.001 * g * g) / (1 - 0.999)
    expect = w0 - lr * mhat / (np.sqrt(vhat) + 1e-8)
    assert np.max(np.abs(w1 - expect)) <= 1e-6
    m.close()


def test_mlp_training_reduces_loss():
    m = Q.MLP(13, 64, 3, 8, 2, seed=11)
    S = 64
    rng = np.random.default_rng(4)
    X = rng.normal(size=(S, m.in_dim)).astype(np.float32)
    As = np.stack([spd_unit_diag(m.P, 500 + s) for s in range(S)]).astype(np.float32)
    m.set_data(X, As)
    idx = np.arange(S, dtype=np.int32)
    first = m.loss_grad(idx)[0]
    for it in range(150):
        m.train_step(rng.choice(S, 16, replace=False).astype(np.int32), 2e-3)
    last = m.loss_grad(idx)[0]
    print(f"Kaporin loss {first:.3f} -> {last:.3f}")
    assert last < first - 0.5
    m.close()


def test_checkpoint_round_trip(tmp_path):
    """A checkpoint restores the network: same weights and bit-identical forward output; with the
    optimizer state, the next Adam steps of the reloaded network equal those of the original."""
    m = Q.MLP(13, 64, 3, 8, 2, seed=21)
    S = 32
    rng = np.random.default_rng(9)
    X = rng.normal(size=(S, m.in_dim)).astype(np.float32)
    As = np.stack([spd_unit_diag(m.P, 700 + s) for s in range(S)]).astype(np.float32)
    m.set_data(X, As)
    for _ in range(5):
        m.train_step(rng.choice(S, 8, replace=False).astype(np.int32), 2e-3)
    f = tmp_path / "pc.mlp"
    m.save(f, with_optimizer=True)
    m2 = Q.MLP.load(f)
    assert (m2.in_dim, m2.hidden, m2.hidden_layers, m2.P, m2.rank) == (13, 64, 3, 8, 2)
    assert np.array_equal(m2.get_params(), m.get_params())
    assert np.array_equal(m2.forward(X), m.forward(X))
    m2.set_data(X, As)
    batches = [rng.choice(S, 8, replace=False).astype(np.int32) for _ in range(3)]
    for b in batches:
        l1 = m.train_step(b, 2e-3)
        l2 = m2.train_step(b, 2e-3)
        assert l1 == l2
    assert np.array_equal(m2.get_params(), m.get_params())
    # weights only: the optimizer restarts
    m.save(f, with_optimizer=False)
    m3 = Q.MLP.load(f)
    assert np.array_equal(m3.forward(X), m.forward(X))
    # a training set and a GPU-computed QGT round-trip through their files too
    Q.dataset_save(tmp_path / "t.set", X, As)
    X4, As4 = Q.dataset_load(tmp_path / "t.set")
    assert np.array_equal(X4, X) and np.array_equal(As4, As)
    th = np.random.default_rng(5).uniform(0, 2 * np.pi, 2 * 6 * 2)
    Qg = Q.qgt(6, Q.hea_ops(6, 2, th))
    Q.qgt_save(tmp_path / "g.qgt", 6, Qg, th)
    n, Qr, thr = Q.qgt_load(tmp_path / "g.qgt")
    assert n == 6 and np.array_equal(Qr, Qg) and np.array_equal(thr, th)
    for x in (m, m2, m3):
        x.close()
