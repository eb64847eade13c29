"""Learned preconditioner workload: natural-gradient VQE on a transverse-field
Ising chain, H = -J sum Z_i Z_{i+1} - h sum X_i, with the hardware-efficient
ansatz, and a small MLP that emits a diagonal + low-rank preconditioner for
CG on (Re Q + lambda I) x = grad E.

Everything runs on the GPU through libqgtl_mi355x: the training pairs
(theta, Q) come from the batched QGT (uniform theta) and from NG-VQE
trajectories (single plans, Cholesky steps); the network is trained by the
library's HIP forward / backward / Adam kernels; the solves are the library's
FP64 GPU solvers, with the exact Cholesky solve as the yardstick of every
iterative solution.

    python -m quantum_geometric_tensor_amd.learned_pc --out result.json [--samples 20000]
"""
import argparse
import json
import time

import numpy as np

from . import (MLP, MODE_MIXED, SOLVE_CG, SOLVE_CHOLESKY, SOLVE_JACOBI_PCG, SOLVE_PCG, Batch, Plan, dataset_save,
               hea_ops, ng_step, solve, tfim_terms)
from .workloads import SplitMix64


def jacobi_scaled(A):
    """D = diag(A) and the unit-diagonal As = D^-1/2 A D^-1/2 (float32) of a stack of matrices."""
    D = np.einsum("sii->si", A)
    s = 1.0 / np.sqrt(D)
    return D, (A * s[:, :, None] * s[:, None, :]).astype(np.float32)


def features(theta, D, h):
    """Same layout as qgtl_pc_features: sin theta, cos theta, log D, field."""
    return np.concatenate([np.sin(theta), np.cos(theta), np.log(D), np.asarray(h)[:, None]], 1).astype(np.float32)


def preconditioner_from_output(Y, D, P, r):
    """d (S x P), U (S x P x r) of M^-1 = diag(d) + U U^T from network outputs (as qgtl_mlp_preconditioner)."""
    Y = Y.reshape(Y.shape[0], P, 1 + r).astype(np.float64)
    d = np.exp(Y[:, :, 0]) / D
    U = Y[:, :, 1:] / np.sqrt(P) / np.sqrt(D)[:, :, None]
    return d, U


def uniform_samples(count, n, L, g, lam, batch=256, log=print):
    """theta ~ U[0, 2pi), h ~ U[0.5, 1.5]; Q from the batched kernel."""
    P = 2 * n * L
    thetas = np.zeros((count, P))
    hs = np.zeros(count)
    for i in range(count):
        hs[i] = 0.5 + g.uniform()
        thetas[i] = g.angles(P)
    A = np.zeros((count, P, P), np.float64)
    b = None
    for c0 in range(0, count, batch):
        c1 = min(count, c0 + batch)
        circuits = [(n, hea_ops(n, L, thetas[i])) for i in range(c0, c1)]
        if b is None or b.count != c1 - c0:
            if b is not None:
                b.close()
            b = Batch(circuits)
        else:
            b.set_angles([c[1]["angle"] for c in circuits])
        Qs, st = b.execute()
        for j, q in enumerate(Qs):
            A[c0 + j] = q.real + lam * np.eye(P)
        log(f"  uniform {c1}/{count}: batch of {c1 - c0} in {st['device_ms']:.1f} ms device")
    if b is not None:
        b.close()
    return thetas, hs, A


def trajectory_samples(count, n, L, g, lam, eta, steps, log=print):
    """NG-VQE trajectories (Cholesky steps) from uniform starts; every step's (theta, Q, grad E) is a sample."""
    P = 2 * n * L
    thetas, hs, A, G, E = [], [], [], [], []
    plan = None
    t0 = time.time()
    while len(thetas) < count:
        h = 0.5 + g.uniform()
        th = g.angles(P)
        if plan is None:
            plan = Plan(n, hea_ops(n, L, th), MODE_MIXED)
        plan.set_hamiltonian(tfim_terms(n, 1.0, h))
        for _ in range(steps):
            if len(thetas) >= count:
                break
            plan.set_angles(hea_ops(n, L, th)["angle"])
            Q, _ = plan.execute()
            e, grad = plan.energy_gradient()
            Ak = Q.real + lam * np.eye(P)
            x, info = solve(Ak, grad, 0.0, SOLVE_CHOLESKY)
            thetas.append(th.copy())
            hs.append(h)
            A.append(Ak)
            G.append(grad)
            E.append(e)
            th = th - eta * x
        log(f"  trajectories: {len(thetas)}/{count} samples, E_last {E[-1]:.4f}, {time.time() - t0:.0f} s")
    plan.close()
    return np.array(thetas), np.array(hs), np.array(A), np.array(G), np.array(E)


def gradients_for(thetas, hs, n, L):
    """grad E at given points (for held-out uniform samples)."""
    P = 2 * n * L
    G = np.zeros((len(thetas), P))
    plan = Plan(n, hea_ops(n, L, thetas[0]), MODE_MIXED)
    for i, (th, h) in enumerate(zip(thetas, hs)):
        plan.set_hamiltonian(tfim_terms(n, 1.0, h))
        plan.set_angles(hea_ops(n, L, th)["angle"])
        plan.execute(want_q=False)
        G[i] = plan.energy_gradient()[1]
    plan.close()
    return G


def train(mlp, X, As, train_idx, test_idx, epochs, batch, lr, g, log=print):
    mlp.set_data(X, As)
    hist = []
    t0 = time.time()
    steps = 0
    for ep in range(epochs):
        order = np.array(train_idx)
        for i in range(len(order) - 1, 0, -1):          # Fisher-Yates with the workload stream
            j = int(g.uniform() * (i + 1))
            order[i], order[j] = order[j], order[i]
        tot = 0.0
        for b0 in range(0, len(order) - batch + 1, batch):
            tot += mlp.train_step(order[b0:b0 + batch].astype(np.int32), lr)
            steps += 1
        nb = max(1, (len(order) - batch + 1 + batch - 1) // batch)
        test = np.mean([mlp.loss_grad(np.asarray(test_idx[c:c + 512], np.int32))[0]
                        for c in range(0, len(test_idx), 512)])
        hist.append({"epoch": ep + 1, "train_loss": tot / nb, "test_loss": float(test)})
        log(f"  epoch {ep + 1}: train {tot / nb:.3f} test {test:.3f} ({time.time() - t0:.1f} s, {steps} steps)")
    return hist, time.time() - t0


def solve_stats(A, G, method, tol, d=None, U=None, maxit=20000):
    x, infos = solve(A, G, 0.0, method, tol=tol, maxit=maxit, d=d, U=U, check=False)
    it = np.array([i["iterations"] for i in infos])
    st = np.array([i["status"] for i in infos])
    return x, it, st, infos[0]["device_ms"]


def run(n=16, L=6, samples=20000, traj_fraction=0.5, lam=1e-3, eta=0.1, traj_steps=250, rank=8, hidden=256,
        hidden_layers=3, epochs=20, batch=256, lr=1e-3, tol=1e-8, seed=0, ng_runs=3, ng_steps=30, log=print,
        save_prefix=None):
    g = SplitMix64(seed)
    P = 2 * n * L
    out = {"config": dict(n=n, L=L, P=P, samples=samples, traj_fraction=traj_fraction, lam=lam, eta=eta,
                          traj_steps=traj_steps, rank=rank, hidden=hidden, hidden_layers=hidden_layers, epochs=epochs,
                          batch=batch, lr=lr, tol=tol, seed=seed)}
    t0 = time.time()
    n_traj = int(round(samples * traj_fraction))
    n_uni = samples - n_traj
    log(f"data: {n_uni} uniform + {n_traj} trajectory samples at n={n} L={L} P={P}")
    th_u, h_u, A_u = uniform_samples(n_uni, n, L, g, lam, log=log)
    th_t, h_t, A_t, G_t, E_t = trajectory_samples(n_traj, n, L, g, lam, eta, traj_steps, log=log)
    thetas = np.concatenate([th_u, th_t])
    hs = np.concatenate([h_u, h_t])
    A = np.concatenate([A_u, A_t])
    del A_u, A_t
    out["data_s"] = time.time() - t0
    D, As = jacobi_scaled(A)
    X = features(thetas, D, hs)
    S = len(thetas)
    perm = np.arange(S)
    for i in range(S - 1, 0, -1):
        j = int(g.uniform() * (i + 1))
        perm[i], perm[j] = perm[j], perm[i]
    n_test = S // 10
    test_idx, train_idx = np.sort(perm[:n_test]), np.sort(perm[n_test:])

    mlp = MLP(3 * P + 1, hidden, hidden_layers, P, rank, seed=seed + 1)
    log(f"training: {len(train_idx)} pairs, {mlp.num_params} parameters")
    hist, train_s = train(mlp, X, As, train_idx, test_idx, epochs, batch, lr, g, log=log)
    out["train"] = {"history": hist, "seconds": train_s, "parameters": int(mlp.num_params)}
    if save_prefix:
        # the training set and the trained network in the library's file formats
        dataset_save(save_prefix + ".set", X, As)
        mlp.save(save_prefix + ".mlp", with_optimizer=True)
        out["saved"] = [save_prefix + ".set", save_prefix + ".mlp"]

    # held-out solves: right-hand sides grad E (trajectory samples have it; uniform ones get it now)
    G = np.zeros((S, P))
    G[n_uni:] = G_t
    uni_test = test_idx[test_idx < n_uni]
    if len(uni_test):
        G[uni_test] = gradients_for(thetas[uni_test], hs[uni_test], n, L)
    At, Gt, Dt = A[test_idx], G[test_idx], D[test_idx]
    Y = mlp.forward(X[test_idx])
    d_l, U_l = preconditioner_from_output(Y, Dt, P, rank)
    x_ch, _, st_ch, ms_ch = solve_stats(At, Gt, SOLVE_CHOLESKY, tol)
    res = {}
    for name, method, d, U in [("cg", SOLVE_CG, None, None), ("jacobi_pcg", SOLVE_JACOBI_PCG, None, None),
                               ("learned_pcg", SOLVE_PCG, d_l, U_l)]:
        x, it, st, ms = solve_stats(At, Gt, method, tol, d, U)
        r = np.linalg.norm(np.einsum("sij,sj->si", At, x) - Gt, axis=1) / np.linalg.norm(Gt, axis=1)
        e = np.linalg.norm(x - x_ch, axis=1) / np.linalg.norm(x_ch, axis=1)
        uni = test_idx < n_uni
        res[name] = {"iterations_mean": float(it.mean()), "iterations_median": float(np.median(it)),
                     "iterations_uniform": float(it[uni].mean()) if uni.any() else None,
                     "iterations_trajectory": float(it[~uni].mean()) if (~uni).any() else None,
                     "converged": int((st == 0).sum()), "systems": int(len(st)),
                     "true_rel_residual_max": float(r.max()), "rel_diff_vs_cholesky_max": float(e.max()),
                     "batch_device_ms": ms}
        log(f"  {name}: iterations mean {it.mean():.1f}, converged {(st == 0).sum()}/{len(st)}, "
            f"max residual {r.max():.2e}, max diff vs Cholesky {e.max():.2e}, batch {ms:.1f} ms")
    res["cholesky"] = {"status_ok": int((st_ch == 0).sum()), "batch_device_ms": ms_ch}
    res["iteration_ratio_jacobi_over_learned"] = res["jacobi_pcg"]["iterations_mean"] / res["learned_pcg"]["iterations_mean"]
    out["held_out"] = res

    # NG-VQE steps/s with each solver, from the same starts
    ng = {}
    starts = [(0.5 + g.uniform(), g.angles(P)) for _ in range(ng_runs)]
    plan = Plan(n, hea_ops(n, L, starts[0][1]), MODE_MIXED)
    for name, method, model in [("jacobi_pcg", SOLVE_JACOBI_PCG, None), ("learned_pcg", SOLVE_PCG, mlp),
                                ("cholesky", SOLVE_CHOLESKY, None)]:
        step_ms, solve_ms, iters, energies, fb = [], [], [], [], 0
        for h, th0 in starts:
            plan.set_hamiltonian(tfim_terms(n, 1.0, h))
            th = th0.copy()
            for _ in range(ng_steps):
                info = ng_step(plan, th, lam, eta, method, tol, 20000, model, h)
                step_ms.append(info["step_ms"])
                solve_ms.append(info["solve_ms"])
                iters.append(info["solve_iterations"])
                fb += info["fallback"]
            energies.append(info["energy"])
        ng[name] = {"steps_per_s": 1e3 / float(np.mean(step_ms[1:])), "step_ms_mean": float(np.mean(step_ms[1:])),
                    "solve_ms_mean": float(np.mean(solve_ms[1:])), "iterations_mean": float(np.mean(iters)),
                    "final_energies": energies, "fallbacks": fb}
        log(f"  NG {name}: {ng[name]['steps_per_s']:.1f} steps/s, solve {ng[name]['solve_ms_mean']:.2f} ms, "
            f"iterations {np.mean(iters):.1f}, fallbacks {fb}")
    plan.close()
    ng["steps_per_s_ratio_learned_over_jacobi"] = ng["learned_pcg"]["steps_per_s"] / ng["jacobi_pcg"]["steps_per_s"]
    ng["solve_ms_ratio_jacobi_over_learned"] = ng["jacobi_pcg"]["solve_ms_mean"] / ng["learned_pcg"]["solve_ms_mean"]
    out["ng_vqe"] = ng
    out["seconds"] = time.time() - t0
    mlp.close()
    return out


def ideal_lowrank_bound(n=16, L=6, samples=400, lam=1e-3, eta=0.1, traj_steps=250, ranks=(8, 16, 32), tol=1e-8,
                        seed=7, log=print):
    """What the preconditioner family can give at best: PCG iterations with M^-1 = D^-1 + U U^T where
    U holds the EXACT eigenvectors of the r smallest eigenvalues of the Jacobi-scaled matrix
    D^-1/2 A D^-1/2 (deflated to 1), against Jacobi-PCG, on the learned workload's systems (half
    uniform theta, half NG trajectories, right-hand side grad E).  A network can only approximate
    these U, so the iteration ratios bound what any diag + rank-r prediction can achieve."""
    g = SplitMix64(seed)
    P = 2 * n * L
    th_u, h_u, A_u = uniform_samples(samples // 2, n, L, g, lam, log=log)
    th_t, h_t, A_t, G_t, _ = trajectory_samples(samples - samples // 2, n, L, g, lam, eta, traj_steps, log=log)
    A = np.concatenate([A_u, A_t])
    G = np.concatenate([gradients_for(th_u, h_u, n, L), G_t])
    D = np.einsum("sii->si", A)
    s = 1.0 / np.sqrt(D)
    As = A * s[:, :, None] * s[:, None, :]
    w, V = np.linalg.eigh(As)
    _, it_j, st_j, _ = solve_stats(A, G, SOLVE_JACOBI_PCG, tol)
    out = {"systems": int(len(A)), "P": P, "jacobi_iterations_mean": float(it_j.mean()),
           "kappa_scaled_median": float(np.median(w[:, -1] / w[:, 0])), "ideal": {}}
    for r in ranks:
        scale = np.sqrt(np.clip(1.0 / w[:, :r] - 1.0, 0.0, None))
        U = s[:, :, None] * V[:, :, :r] * scale[:, None, :]
        _, it, st, _ = solve_stats(A, G, SOLVE_PCG, tol, 1.0 / D, U)
        out["ideal"][r] = {"iterations_mean": float(it.mean()), "converged": int((st == 0).sum()),
                           "iteration_ratio_jacobi_over_ideal": float(it_j.mean() / it.mean())}
        log(f"rank {r}: ideal deflation {it.mean():.1f} iterations vs Jacobi {it_j.mean():.1f} "
            f"({it_j.mean() / it.mean():.2f}x), converged {(st == 0).sum()}/{len(st)}")
    return out


def pcg_iteration_cost(sizes=(64, 192, 240, 512), rank=8, log=print):
    """Device time per PCG iteration of the library's FP64 solver (one system, one workgroup):
    the difference between 550 and 50 iterations (tolerance 1e-300, never met) over 500, best
    of three, for Jacobi-PCG and diagonal + rank-`rank` PCG on an SPD matrix with spectrum
    1e-7 .. 1."""
    out = {}
    for P in sizes:
        rng = np.random.default_rng(P)
        Qm, _ = np.linalg.qr(rng.normal(size=(P, P)))
        A = (Qm * np.logspace(-7, 0, P)) @ Qm.T
        A = 0.5 * (A + A.T)
        b = rng.normal(size=P)
        d = 1.0 / np.diag(A)
        U = Qm[:, :rank] * 0.01
        row = {}
        for name, method, dd, UU in (("jacobi_pcg", SOLVE_JACOBI_PCG, None, None), ("lowrank_pcg", SOLVE_PCG, d, U)):
            t = {}
            for maxit in (50, 550):
                t[maxit] = min(solve(A, b, 0.0, method, tol=1e-300, maxit=maxit, d=dd, U=UU, check=False)[1]["device_ms"]
                               for _ in range(3))
            row[name] = (t[550] - t[50]) / 500 * 1e3
        out[P] = row
        log(f"P={P}: Jacobi-PCG {row['jacobi_pcg']:.2f} us/iteration, diagonal + rank-{rank} "
            f"{row['lowrank_pcg']:.2f} us/iteration")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pcg-timing", action="store_true", help="only time PCG iterations and exit")
    ap.add_argument("--ideal-bound", action="store_true",
                    help="only measure PCG with exact smallest-eigenvector deflation (the family's best case) and exit")
    ap.add_argument("--out", default=None)
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--layers", type=int, default=6)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--rank", type=int, default=8)
    ap.add_argument("--traj-steps", type=int, default=250)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--save", default=None, help="path prefix: write <prefix>.set (training set) and <prefix>.mlp")
    args = ap.parse_args()
    if args.pcg_timing:
        pcg_iteration_cost(rank=args.rank, log=lambda s: print(s, flush=True))
        return
    if args.ideal_bound:
        res = ideal_lowrank_bound(n=args.n, L=args.layers, traj_steps=args.traj_steps,
                                  log=lambda s: print(s, flush=True))
        print(json.dumps(res, indent=1), flush=True)
        return
    res = run(n=args.n, L=args.layers, samples=args.samples, epochs=args.epochs, rank=args.rank,
              traj_steps=args.traj_steps, lr=args.lr, log=lambda s: print(s, flush=True), save_prefix=args.save)
    text = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text)
    print(json.dumps({k: res[k] for k in ("held_out", "ng_vqe")}, indent=1), flush=True)


if __name__ == "__main__":
    main()
