"""Benchmark of the Schmidt-form MPS update path on one MI355X.

Workload: one trajectory of N=64 qubits, Haar brickwork circuit with
projective Z measurement of each qubit with probability p=0.1 after every
layer, Const(chi_max=256) truncation, complex128 throughout.  A "step" is one
circuit layer (the brickwork gates plus the measurement layer with its
refactor sweeps) once the bond dimension has saturated; the warmup layers
grow the state from |0...0> to saturation.  The metric is two-site updates
per second (contract -> SVD -> truncate -> split, gates and refactors alike).

Prints one JSON line (stdout); progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
for _p in (REPO, REPO / "tensor-net_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

METRIC = "gate+SVD updates/s @chi=256 FP64 vs mixed; trajectories/s + p99 latency"
# HBM bytes per dispatch of the dominant Jacobi kernels on this workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a 1 GiB copy calibration
# (profiles/jacobi_single_traj_chi256_hbm_traffic.md).
TRAFFIC_BYTES = {"k_jacobi_apply<8>": 251.3 * 1024, "k_jacobi_gram<8>": 82.7 * 1024}
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(st, n, trunc, layer, budget_s):
    """Time the CPU oracle (numpy + LAPACK zgesvd, one thread) on the saturated
    state: the gates of the next layer, then refactors, until the budget."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle.mps_oracle import VidalMPS

    o = VidalMPS(n, trunc)
    o.G = [np.ascontiguousarray(st.gamma(k)) for k in range(n)]
    o.L = [st.schmidt(b) for b in range(n - 1)]
    count = 0
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        unis, _ = layer
        for k, u in unis:
            o.apply_unitary2(k, u)
            count += 1
            if time.perf_counter() - t0 > budget_s:
                break
        j = 0
        while time.perf_counter() - t0 <= budget_s:
            o.local_refactor(j % (n - 1))
            j += 1
            count += 1
        el = time.perf_counter() - t0
    return count / el, count, el


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--chi", type=int, default=256)
    ap.add_argument("--p", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=10546)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        nbytes = args.calibrate_copy << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 3}), flush=True)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the MPS path shards by trajectory and this build "
              "runs on a single device (use --gpus 1)", file=sys.stderr)
        return 2

    from tnmps import MPS
    from tnmps.workload import haar_circuit

    n, chi, p = args.n, args.chi, args.p
    W, K = args.warmup, args.steps
    trunc = ("const", chi)
    c = haar_circuit(n, W + K + 1, p, args.seed)
    draws = iter(c.draws)
    st = MPS.new_qubits(n, trunc)
    t_w = time.perf_counter()
    for t in range(W):
        st.run_circuit(c.layers[t:t + 1], rs=draws)
        log(f"warmup layer {t + 1}/{W}: max bond {st.max_bond_dim()} "
            f"({time.perf_counter() - t_w:.1f}s)")
    st.sync()
    st.set_timing(True)
    st.reset_perf()
    u0, _ = st.stats()
    t0 = time.perf_counter()
    for t in range(W, W + K):
        st.run_circuit(c.layers[t:t + 1], rs=draws)
        log(f"timed layer {t - W + 1}/{K}: max bond {st.max_bond_dim()} "
            f"({time.perf_counter() - t0:.1f}s)")
    st.sync()
    t1 = time.perf_counter()
    st.set_timing(False)
    u1, _ = st.stats()
    perf = st.perf()
    updates = u1 - u0
    wall = t1 - t0
    value = updates / wall

    # dominant kernel: the Jacobi step kernel, credited with the algorithmic
    # flops of the SVDs it performs
    jac_s = perf["ms_jacobi"] / 1e3
    achieved = perf["svd_flops"] / jac_s / 1e12 if jac_s > 0 else 0.0
    # the Jacobi step kernels (row-split gram + solve/apply, or one-workgroup
    # steps) are credited with the algorithmic R-SVD flops of the SVDs they
    # perform, over the HIP-event span of the Jacobi sweeps on the library's
    # stream (dispatch gaps between dependent launches included)
    roof = {
        "bound": "mfma",
        "kernel": "jacobi steps: k_jacobi_apply<8> + k_jacobi_gram<8> (+ k_jacobi_step<8>)",
        "achieved": round(achieved, 4),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
        "traffic": round((TRAFFIC_BYTES["k_jacobi_apply<8>"] + TRAFFIC_BYTES["k_jacobi_gram<8>"]) / 2),
        "avg_launch_us": round(perf["ms_jacobi"] * 1e3 / max(perf["jacobi_launches"], 1), 3),
        "launches": perf["jacobi_launches"],
        "svd_flops": perf["svd_flops"],
    }
    breakdown = {k: round(perf[k], 3) for k in ("ms_theta", "ms_jacobi", "ms_post", "ms_measure",
                                                "ms_host_launch")}
    breakdown["sweeps_per_update_batch"] = round(perf["sweeps"] / max(perf["updates"], 1), 3)
    breakdown["alg_tflops_whole_update"] = round(perf["alg_flops"] / wall / 1e12, 4)

    cpu = None
    if not args.no_cpu:
        rate, cnt, el = cpu_baseline(st, n, trunc, c.layers[W + K], args.cpu_budget)
        cpu = {"value": round(rate, 4), "unit": "two-site updates/s", "cores": 1,
               "kind": "port",
               "sample": f"oracle (numpy + LAPACK zgesvd, 1 thread) on the saturated "
                         f"N={n} chi={chi} state: {cnt} updates (next layer's gates, then "
                         f"refactors) in {el:.1f}s"}
        log(f"cpu baseline: {rate:.3f} updates/s over {cnt} updates")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "two-site updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"single trajectory N={n} Haar brickwork, p={p} Z measurements, "
                        f"Const(chi_max={chi}), complex128, plain FP64 GPU path; "
                        f"step = one saturated circuit layer",
            "n_qubits": n, "p": p, "chi_max": chi, "trajectories": 1,
            "updates_timed": updates,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "breakdown": breakdown,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
