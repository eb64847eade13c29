"""Headline benchmark: QGT/s at n=20, L=6 (P=240) on one MI355X.

Workload (the precision form of the QGT path): hardware-efficient ansatz with
n=20 qubits and 6 layers (RY, RZ on every qubit, then a CNOT ring), P=240
parameters drawn uniform in [0, 2pi) from a splitmix64 stream (seed 0).  One
step = one full P x P quantum geometric tensor through the mixed-precision HIP
path (complex64 sweeps, split-BF16 MFMA Gram with FP32 accumulation, FP64
reductions and correction), including the download of Q.  Reported beside it:
the plain FP64 GPU path (complex128, one kernel per gate, FP64 Gram), the max
normwise relative error of the mixed result against both FP64 results, the
roofline fraction of the dominant kernel (the LDS-tiled sweep pass), and the
FP64 CPU oracle timed on this host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
N_QUBITS, LAYERS = 20, 6


def splitmix_angles(seed, count):
    """Same stream as the oracle's oracle_uniform_angles (splitmix64, (x >> 11) 2^-53)."""
    out = np.empty(count, np.float64)
    s = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    for i in range(count):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[i] = 2.0 * np.pi * ((z >> 11) * (1.0 / 9007199254740992.0))
    return out


TUNE_FILE = os.path.join(ROOT, "quantum_geometric_tensor_amd", "tune", "batch_hea_n12-18_L2-8.json")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_pass_regs_hea_n20_L6.txt")


def read_traffic(kernel_prefix):
    """HBM bytes per dispatch of the dominant kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated with a plain copy), or None."""
    try:
        vals = {}
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if ":" in line:
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
        if not vals.get("kernel", "").startswith(kernel_prefix):
            return None
        return float(vals["hbm_bytes_per_dispatch"])
    except (OSError, KeyError, ValueError):
        return None


def batched_leg(Q, reps=10, warmup=2):
    """512 independent HEA instances (n ~ U{12..18}, L ~ U{2..8}, then theta, one
    splitmix64 seed-0 stream) in one captured launch sequence; each timed run
    sets the angles (host plan rebuilds, descriptor upload) and runs the batch,
    once with the Q matrices left in HBM (value) and once downloading them all."""
    from quantum_geometric_tensor_amd.workloads import hea_batch
    work = hea_batch(512, seed=0)
    tune = Q.load_batch_tune(TUNE_FILE) if os.path.exists(TUNE_FILE) else None
    b = Q.Batch([(n, Q.hea_ops(n, L, th)) for n, L, th in work], tune)
    angles = [Q.hea_ops(n, L, th)["angle"] for n, L, th in work]
    qbuf = np.empty(b.q_size, np.complex128)       # the caller's result buffer, reused per execution
    for _ in range(warmup):
        b.set_angles(angles)
        b.execute(out=qbuf)
    # timed: new angles (host plan rebuilds, descriptor upload) and the launch sequence, the 512 Q
    # matrices left in HBM; then the same with every Q copied to the host (PCIe)
    walls, devs, ups, sets = [], [], [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        b.set_angles(angles)
        t1 = time.perf_counter()
        _, st = b.execute(want_q=False)
        walls.append(time.perf_counter() - t0)
        sets.append(t1 - t0)
        devs.append(st["device_ms"])
        ups.append(st["upload_ms"])
    walls_q, downs = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        b.set_angles(angles)
        _, st = b.execute(out=qbuf)
        walls_q.append(time.perf_counter() - t0)
        downs.append(st["download_ms"])
    wall = float(np.median(walls))
    wall_q = float(np.median(walls_q))
    out = {"workload": "512 HEA instances, n~U{12..18}, L~U{2..8}, seed 0", "value": 512 / wall, "unit": "QGT/s",
           "ms_per_batch": wall * 1e3, "device_ms_median": float(np.median(devs)),
           "host_ms_median": {"set_angles": float(np.median(sets)) * 1e3, "upload": float(np.median(ups)),
                              "download": float(np.median(downs))},
           "with_q_download": {"value": 512 / wall_q, "unit": "QGT/s", "ms_per_batch": wall_q * 1e3},
           "launches": st["launches"],
           "graph": bool(st["graph"]), "sweep_bytes": st["sweep_bytes"], "gram_flops": st["gram_flops"],
           "device_GB": b.device_bytes / 1e9, "tune": b.get_tune().as_dict(),
           "tune_file": os.path.relpath(TUNE_FILE, ROOT) if tune is not None else None}
    b.close()
    return out


def large_leg(Q, reps=2):
    """Large single instance: n=26, L=8, P=416 (complex64 states of 512 MB; the 416
    derivative states, 223 GB, stay resident in HBM), theta ~ U[0,2pi) seed 0."""
    n, L = 26, 8
    theta = splitmix_angles(0, 2 * n * L)
    plan = Q.Plan(n, Q.hea_ops(n, L, theta), Q.MODE_MIXED)
    t0 = time.perf_counter()
    plan.execute()
    first = time.perf_counter() - t0
    walls = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.execute()
        walls.append(time.perf_counter() - t0)
    Q.set_profiling(True)
    _, st = plan.execute()
    Q.set_profiling(False)
    plan.close()
    wall = float(np.median(walls))
    return {"workload": "HEA QGT n=26 L=8 P=416", "value": 1.0 / wall, "unit": "QGT/s", "ms_per_qgt": wall * 1e3,
            "first_call_ms": first * 1e3,
            "breakdown_ms": {"sweep": st["sweep_ms"], "inject": st["inject_ms"], "gram": st["gram_ms"],
                             "launches": st["kernel_launches"]},
            "sweep_GBps": st["sweep_bytes"] / (st["sweep_ms"] * 1e-3) / 1e9 if st["sweep_ms"] else None,
            "gram_tflops": st["gram_flops"] / (st["gram_kernel_ms"] * 1e-3) / 1e12 if st["gram_kernel_ms"] else None}


def _cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)


def cpu_leg(Q, n, L, theta, Qmix, full=False):
    """The FP64 C oracle as the CPU baseline, in the reference's loop structure:
    derivative states materialised once (exact, complex128), then the per-element
    loop over mu <= nu with three length-2^n dot products per element
    (quantum_geometric_curvature.c:300-317 over quantum_geometric_tensor_network.c:1127-1166)
    and the P x P loop (differential_geometry.c:2832-2858).  The n=20 loops are timed on a
    bounded sample of rows (1 warm-up + 5 repetitions, median) and scaled by element
    count; the n=10 instance is timed whole (median of 5).  full=True adds the
    single-thread n=20 derivative states (minutes) and the n=20 P x P loop."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = _cpu_threads()
    P = 2 * n * L
    total_el = {"elementwise": P * (P + 1) // 2, "pairwise": P * P}
    oops = O.hea_ops(n, L, theta)
    psi, D, deriv_all = O.derivative_states_timed(n, oops, threads)

    def sampled(variant, nthreads, rows, reps=6):
        t, el = O.qgt_loop_time(psi, D, variant, nthreads, rows, reps)
        med = float(np.median(t[1:]))
        return med, el, med * total_el[variant] / el

    rows = min(P, threads)
    med, el, loop_all = sampled("elementwise", threads, rows)
    per_qgt = deriv_all + loop_all
    out = {"value": 1.0 / per_qgt, "unit": "QGT/s", "cores": threads, "kind": "port",
           "sample": (f"n={n} L={L} P={P}, FP64 C oracle on {threads} threads: derivative states once "
                      f"({deriv_all:.2f} s) + the per-element loop mu<=nu with 3 dot products per element, rows "
                      f"0..{rows - 1} ({el} of {total_el['elementwise']} elements) {med:.2f} s (median of 5) scaled "
                      f"by element count: {per_qgt:.1f} s per QGT"),
           "seconds_per_qgt": per_qgt, "derivative_states_s": deriv_all, "loop_s_scaled": loop_all}
    med1, el1, loop1 = sampled("elementwise", 1, 1, reps=4)
    single = {"threads": 1, "loop_s_scaled": loop1, "sample": f"row 0 ({el1} elements) {med1:.2f} s, median of 3"}
    if full:
        _, _, d1 = O.derivative_states_timed(n, oops, 1, psi, D)
        single["derivative_states_s"] = d1
        single["seconds_per_qgt"] = d1 + loop1
        medp, elp, loopp = sampled("pairwise", threads, rows)
        out["pairwise_all_threads"] = {"seconds_per_qgt": deriv_all + loopp, "loop_s_scaled": loopp,
                                       "sample": f"rows 0..{rows - 1} ({elp} elements) {medp:.2f} s, median of 5"}
    out["elementwise_single_thread"] = single
    del D
    # the n=10, L=4 instance whole: both loop forms, one thread and all threads
    n1, L1 = 10, 4
    ops1 = O.hea_ops(n1, L1, O.uniform_angles(0, 2 * n1 * L1))
    P1 = 2 * n1 * L1
    small = {}
    for nt in (1, threads):
        ds = []
        for _ in range(6):
            psi1, D1, d = O.derivative_states_timed(n1, ops1, nt)
            ds.append(d)
        for variant in ("elementwise", "pairwise"):
            t, _ = O.qgt_loop_time(psi1, D1, variant, nt, P1, 6)
            small[f"{variant}_{nt}_threads_s"] = float(np.median(ds[1:])) + float(np.median(t[1:]))
    out["n10_L4_whole"] = small
    # accuracy of the mixed GPU result against the oracle's full matrix (hoisted form, same values)
    Qcpu = O.qgt_hoisted(n, oops)
    return out, Q.rel_err(Qmix, Qcpu)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline with the single-thread n=20 derivative states and the n=20 P x P loop (minutes)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time only the mixed-precision path (for profiler runs): no FP64 legs")
    ap.add_argument("--no-batched", action="store_true", help="skip the 512-instance batched leg")
    ap.add_argument("--batched-only", action="store_true", help="run only the batched leg (for profiler runs)")
    ap.add_argument("--large", action="store_true",
                    help="also time the large single instance n=26, L=8 (P=416, ~225 GB of device memory)")
    ap.add_argument("--large-only", action="store_true", help="run only the large single instance")
    ap.add_argument("--one-stream", action="store_true",
                    help="per-launch timing on for every run, which keeps all kernels on the library stream "
                         "(so a profiler sees the launches the HIP-event breakdown times)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one plain copy kernel of this size (traffic-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the QGT path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration).", file=sys.stderr)
        sys.exit(2)

    import quantum_geometric_tensor_amd as Q
    if args.copy_calibration_mib:
        import ctypes
        ms = ctypes.c_double(0.0)
        nbytes = args.copy_calibration_mib << 20
        rc = Q.lib().qgtl_debug_copy_kernel(ctypes.c_size_t(nbytes), ctypes.byref(ms))
        if rc:
            raise Q.QGTLError(rc, Q.lib().qgtl_last_error().decode())
        print(json.dumps({"copy_calibration_bytes": nbytes, "ms": ms.value,
                          "GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)
    if args.batched_only:
        print(json.dumps(batched_leg(Q)), flush=True)
        return
    if args.large_only:
        print(json.dumps(large_leg(Q)), flush=True)
        return
    n, L = N_QUBITS, LAYERS
    theta = splitmix_angles(0, 2 * n * L)
    ops = Q.hea_ops(n, L, theta)
    P = Q.num_params(ops)
    N = 1 << n

    plan = Q.Plan(n, ops, Q.MODE_MIXED)
    Qmix = np.empty((P, P), np.complex128)
    if args.one_stream:
        Q.set_profiling(True)
    for _ in range(args.warmup):
        plan.execute(out=Qmix)
    # the rate with Q brought back to the host (PCIe) each step, then the timed steps of `value`:
    # Q stays in HBM (each call still waits for its GPU work)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.execute(out=Qmix)
    t1 = time.perf_counter()
    ms_per_step_download = (t1 - t0) * 1e3 / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, st_last = plan.execute(want_q=False)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    qgt_per_s = 1e3 / ms_per_step
    # per-call protocol: the median over 20 calls of each call's own time (descriptor upload,
    # the launch sequence -- a replayed graph once the plan is unchanged)
    per_call = []
    for _ in range(20):
        _, stc = plan.execute(want_q=False)
        per_call.append(stc["wall_ms"])
    per_call_median = float(np.median(per_call))

    # per-kernel device times (HIP events on the library's stream) in a separate pass
    Q.set_profiling(True)
    _, st = plan.execute()
    _, st = plan.execute()
    Q.set_profiling(bool(args.one_stream))
    sweep_avg_ms = st["sweep_ms"] / max(1, st["sweep_launches"])
    sweep_bytes_per_launch = st["sweep_bytes"] / max(1, st["sweep_launches"])
    achieved = sweep_bytes_per_launch / (sweep_avg_ms * 1e-3) / 1e9
    plan.close()
    if args.mixed_only:
        print(json.dumps({"mixed_only": True, "value": qgt_per_s, "unit": "QGT/s", "ms_per_step": ms_per_step,
                          "breakdown_ms": {"sweep": st["sweep_ms"], "inject": st["inject_ms"],
                                           "gram": st["gram_ms"], "launches": st["kernel_launches"]},
                          "sweep_achieved_GBps": achieved}), flush=True)
        return

    # plain FP64 GPU path (the speed-up denominator and the GPU FP64 cross-check)
    plain = Q.Plan(n, ops, Q.MODE_FP64_PLAIN)
    Q64, _ = plain.execute()
    tp0 = time.perf_counter()
    for _ in range(args.plain_steps):
        Q64, _ = plain.execute()
    tp1 = time.perf_counter()
    plain.close()
    plain_ms = (tp1 - tp0) * 1e3 / args.plain_steps
    err_plain = Q.rel_err(Qmix, Q64)

    cpu = None
    err_cpu = None
    if not args.no_cpu_baseline:
        cpu, err_cpu = cpu_leg(Q, n, L, theta, Qmix, full=args.cpu_full)

    line = {
        "metric": "QGT/s at n=20,P=240: mixed-precision vs FP64 plain on GPU, rel err <=1e-6",
        "value": qgt_per_s,
        "unit": "QGT/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 sweeps + split-fp16x3 MFMA Gram (fp32 acc) + f64 reductions/correction",
        "data": "synthetic (theta ~ U[0,2pi), splitmix64 seed 0)",
        "config": {"workload": "hardware-efficient ansatz QGT, n=20 qubits, 6 layers, P=240", "n": n,
                   "layers": L, "P": P, "amplitudes": N, "mode": "mixed"},
        "fp64_plain_gpu": {"value": 1e3 / plain_ms, "unit": "QGT/s", "ms_per_qgt": plain_ms},
        "per_call_median_ms": per_call_median,
        "with_q_download": {"value": 1e3 / ms_per_step_download, "unit": "QGT/s", "ms_per_step": ms_per_step_download,
                            "note": "each step also copies the P x P complex128 Q to host memory over PCIe"},
        "graph_replay": bool(st_last["graph"]),
        "speedup_vs_fp64_plain": plain_ms / ms_per_step,
        "rel_err_vs_fp64_plain_gpu": err_plain,
        "rel_err_vs_fp64_cpu_oracle": err_cpu,
        "roofline": {"bound": "hbm",
                     "kernel": "pass_c64_regs_kernel + pass_c64_inject_kernel (register-phase sweep passes)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("pass_c64"),
                     "bytes_per_launch": sweep_bytes_per_launch, "avg_launch_ms": sweep_avg_ms},
        "breakdown_ms": {"sweep": st["sweep_ms"], "inject": st["inject_ms"], "gram": st["gram_ms"],
                         "finalize": st["finalize_ms"], "launches": st["kernel_launches"]},
        "gram": {"kernel_ms": st["gram_kernel_ms"], "flops": st["gram_flops"],
                 "tflops_algorithmic": st["gram_flops"] / (st["gram_kernel_ms"] * 1e-3) / 1e12 if st["gram_kernel_ms"] else None},
        "cpu_baseline": cpu,
    }
    if not args.no_batched:
        line["batched"] = batched_leg(Q)
    if args.large:
        line["large"] = large_leg(Q)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
