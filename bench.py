"""Benchmark of the Schmidt-form MPS update path on one MI355X.

Workload: one trajectory of N=64 qubits, Haar brickwork circuit with
projective Z measurement of each qubit with probability p=0.1 after every
layer, Const(chi_max=256) truncation, complex128 throughout.  A "step" is one
circuit layer (the brickwork gates plus the measurement layer with its
refactor sweeps) once the bond dimension has saturated: the warmup runs at
least --warmup layers and then until half the bonds that can reach chi_max
are at chi_max (saturated()).  The metric is two-site updates per second
(contract -> SVD -> truncate -> split, gates and refactors alike).

The same run also measures, under "mixed_batched", the metric's other side:
the batched path at the same N, chi and p (256 trajectories in 4 lock-step
groups), with the FP32 contraction (saturated the same way) and then in FP64
throughout on the same circuits and number of warmup layers; each rate's
ratio to the FP64 single-trajectory value of this run, and the FP32 run's
entropy error against an FP64 replay of its first trajectories' outcome
records.  The CPU
baseline runs the oracle in one process per host core (up to 16), each on
its own copy of the saturated state, in a pool forked before the GPU is
touched.

Other workloads (``--workload``), each printing its own JSON line:
  batched  1024 trajectories of N=32, depth 32, p on an 11-point grid
           0.05..0.30, Const(128), dealt into 4 groups advanced in lock step,
           each on its own engine (stream) and host thread; a window of
           layers after warmup is timed.
  mixed    N=48 trajectories with Const(512): the FP32-contraction path
           against the FP64 path on the same outcome record, per-layer
           Schmidt-value and entropy error and both update rates.
  serving  Poisson arrivals of trajectory requests (N in {16,32,64,128},
           p ~ U[0, 0.4], chi_max in {64,128,256}, depth N, FP32 contraction)
           served by continuous batching; trajectories/s and p50/p99 latency.

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
# Memory traffic per dispatch of the dominant Jacobi kernels on this workload:
# rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per dispatch, one counter per run)
# scaled by what the same counters read for a plain 1 GiB device copy
# (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); profiles/jacobi_single_traj_chi256_hbm_traffic.md.
PMC_KIB = {"k_jacobi_fused_r<4, 1>": (1160.3, 621.6)}
PMC_SCALE = (2.0, 1.0)
TRAFFIC_BYTES = {k: 1024.0 * (f * PMC_SCALE[0] + w * PMC_SCALE[1]) for k, (f, w) in PMC_KIB.items()}
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the Jacobi kernels on this workload, "
                  "profiles/jacobi_single_traj_chi256_hbm_traffic.md (not measured in this run)")
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_info() -> dict:
    """Host cores this process may use and the CPU model (/proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(),
            "model": model}


def start_cpu_pool(workers: int):
    """Worker processes for the CPU baseline, forked before anything in this
    process touches the GPU (no HIP state is inherited; the workers only run
    the oracle).  One process per core up to 16, the host share of one GPU."""
    import multiprocessing as mp

    n = workers or min(16, len(os.sched_getaffinity(0)))
    return mp.get_context("fork").Pool(n), n


def _cpu_worker(path: str, n: int, trunc, gates, budget_s: float):
    """One core: the oracle (numpy + LAPACK zgesvd, one thread) on its own
    copy of the saturated state: the next layer's gates, then refactors,
    until the budget.  Returns (updates, seconds)."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle.mps_oracle import VidalMPS

    d = np.load(path)
    o = VidalMPS(n, trunc)
    o.G = [np.ascontiguousarray(d[f"g{k}"]) for k in range(n)]
    o.L = [d[f"l{b}"] for b in range(n - 1)]
    count = 0
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for k, u in gates:
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
    return count, el


def cpu_baseline(pool, workers: int, st, n, trunc, layer, budget_s):
    """All workers at once, each on a copy of the saturated state; the rate is
    the sum of the workers' update rates (one trajectory per core, as the
    reference runs one MPS per rayon thread)."""
    import numpy as np

    path = f"/tmp/tnmps_cpu_state_{os.getpid()}.npz"
    arrs = {f"g{k}": st.gamma(k) for k in range(n)}
    arrs.update({f"l{b}": st.schmidt(b) for b in range(n - 1)})
    np.savez(path, **arrs)
    try:
        unis, _ = layer
        res = pool.starmap(_cpu_worker, [(path, n, trunc, list(unis), budget_s)] * workers)
    finally:
        os.unlink(path)
    rate = sum(c / e for c, e in res)
    return rate, sum(c for c, _ in res), max(e for _, e in res)


def saturated(states, chi: int, frac: float = 0.5) -> bool:
    """True when, in every listed state, at least `frac` of the bonds whose
    exact dimension min(2^(b+1), 2^(n-b-1)) reaches chi are at chi."""
    for st in states:
        n = st.n
        d = st.bond_dims()
        can = [b for b in range(n - 1)
               if min(b + 1, n - b - 1) >= 40 or min(2 ** (b + 1), 2 ** (n - b - 1)) >= chi]
        if not can or sum(1 for b in can if d[b] == chi) < frac * len(can):
            return False
    return True


def run_batched(args) -> dict:
    """Lock-step batch: per-trajectory p from an 11-point grid, uneven bond
    dimensions across the batch; updates/s over the timed layers.  With
    ``--groups G`` the trajectories are dealt round-robin into G lock-step
    groups, each on its own engine (stream) and host thread, so that one
    group's last sweeps (a few slow SVDs) overlap the others' work."""
    import threading

    import numpy as np

    from tnmps import Engine
    from tnmps.workload import haar_circuit_fast

    n, depth, chi, T = args.n or 32, args.depth or 32, args.chi or 128, args.traj or 1024
    W, K = args.warmup, args.steps
    G = max(1, min(args.groups, T))
    lo, hi, npts = args.pgrid.split(",")
    pgrid = np.linspace(float(lo), float(hi), int(npts))
    t_gen = time.perf_counter()
    circs = [haar_circuit_fast(n, depth, float(pgrid[i % len(pgrid)]), args.seed + 7 * i)
             for i in range(T)]
    log(f"generated {T} circuits in {time.perf_counter() - t_gen:.1f}s")
    groups = [list(range(g, T, G)) for g in range(G)]
    engs = [Engine(chi) for _ in range(G)]
    tune_label, tune_rows = tuning_for(args, n, chi, args.prec, batched=True, groups=G)
    if tune_rows is not None:
        if [dict(r) for r in tune_rows] == engs[0].get_tuning():
            tune_label = f"default (the tuned table {tune_label} holds the defaults)"
        for e in engs:
            e.set_tuning(tune_rows)
    states = [[engs[g].new_qubits(n, ("const", chi)) for _ in idx] for g, idx in enumerate(groups)]
    if args.prec != "f64":
        for grp in states:
            for st in grp:
                st.set_precision(args.prec)
    draws = [[iter(circs[i].draws) for i in idx] for idx in groups]
    ents = [[0.0] * len(idx) for idx in groups]
    errors = []
    # outcome records and entropies of the first trajectories of group 0, for
    # the FP64 replay that states a reduced-precision run's error
    n_rep = min(2, len(groups[0])) if args.prec != "f64" else 0
    rep_rec = [[] for _ in range(n_rep)]
    rep_ent = [[] for _ in range(n_rep)]

    def layers(g, t0, t1, label):
        try:
            for t in range(t0, t1):
                res = engs[g].batch_bilayer(states[g], [circs[i].layers[t] for i in groups[g]],
                                            draws=draws[g])
                ents[g] = [r[2] for r in res]
                if g == 0:
                    for j in range(n_rep):
                        rep_rec[j].append(res[j][0])
                        rep_ent[j].append(res[j][2])
                if g == 0:
                    log(f"{label} layer {t + 1} ({time.perf_counter() - t_w:.1f}s)")
        except Exception as exc:  # surfaced after the join
            errors.append(exc)

    def run_all(t0, t1, label):
        th = [threading.Thread(target=layers, args=(g, t0, t1, label)) for g in range(G)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        if errors:
            raise errors[0]
        for e in engs:
            e.sync()

    # warmup: at least W layers, then until the first trajectory of every
    # group is saturated (or the circuit leaves only K layers)
    t_w = time.perf_counter()
    W_run = 0
    while W_run < depth - K and (W_run < W or not saturated([s[0] for s in states], chi)):
        run_all(W_run, W_run + 1, "warmup")
        W_run += 1
    sat = saturated([s[0] for s in states], chi)
    for e in engs:
        e.set_timing(True)
        e.reset_perf()
    u0 = sum(e.stats()[0] for e in engs)
    t0 = time.perf_counter()
    t_w = t0
    run_all(W_run, W_run + K, "timed")
    wall = time.perf_counter() - t0
    for e in engs:
        e.set_timing(False)
    u1 = sum(e.stats()[0] for e in engs)
    perfs = [e.perf() for e in engs]
    perf = {k: sum(pf[k] for pf in perfs) for k in perfs[0]}
    updates = u1 - u0
    # stream time of the Jacobi sweeps: the groups' streams overlap, so the
    # busiest stream bounds the sweeps' share of the wall time
    jac_s = max(pf["ms_jacobi"] for pf in perfs) / 1e3
    achieved = perf["svd_flops"] / wall / 1e12
    mean_ent = float(np.mean([x for e in ents for x in e]))
    replay = None
    if n_rep:
        # the same trajectories in FP64 on their recorded outcomes (untimed)
        from tnmps import MPS
        err = 0.0
        for j in range(n_rep):
            i = groups[0][j]
            ref = MPS.new_qubits(n, ("const", chi))
            got = ref.run_circuit(circs[i].layers[:W_run + K], forced=rep_rec[j])
            err = max(err, max(abs(a - b) for a, b in zip(rep_ent[j], got["entropy"][1:])))
            del ref
        replay = {"trajectories": n_rep, "layers": W_run + K,
                  "max_half_chain_entropy_abs_err_vs_fp64": float(f"{err:.3e}")}
        log(f"FP64 replay of {n_rep} trajectories: entropy error {err:.3e}")
    return {
        "metric": "two-site updates/s (batched lock-step trajectories)",
        "value": round(updates / wall, 2), "unit": "two-site updates/s", "n_gpus": 1,
        "steps": K, "warmup": W_run, "ms_per_step": round(wall / K * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.prec == "f64" else f"{args.prec} contraction, f64 SVD",
        "data": "synthetic",
        "config": {"workload": f"{T} trajectories N={n} depth={depth} p in linspace({lo},{hi},{npts}) "
                               f"Const({chi}), {G} lock-step group(s) on {G} stream(s); "
                               f"step = one layer of every trajectory"
                               + (" after saturation" if sat else " (not saturated)"),
                   "trajectories": T, "groups": G, "n_qubits": n, "chi_max": chi,
                   "saturated": sat, "warmup_layers_min": W,
                   "updates_timed": updates,
                   "trajectory_layers_per_s": round(T * K / wall, 2),
                   "trajectories_per_s_equiv": round(T * K / wall / depth, 3),
                   "mean_half_chain_entropy_last_layer": round(mean_ent, 6),
                   "tuning": tune_label, "fp64_replay": replay},
        "roofline": {"bound": "mfma", "kernel": "jacobi step kernels",
                     "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": None,
                     "flops": "LAPACK-style count of the SVDs posed, over the wall time",
                     "alg_tflops_whole_update": round(perf["alg_flops"] / wall / 1e12, 4),
                     "busiest_stream_jacobi_s": round(jac_s, 3)},
        "breakdown": {k: round(perf[k], 2) for k in ("ms_theta", "ms_qr", "ms_jacobi", "ms_post",
                                                     "ms_measure", "ms_host_launch", "qr_tasks")},
        "memory_kernels": memory_kernels(perf, G),
        "contraction_kernel": contraction_kernel(perf, args.prec, G),
    }


TUNED_TABLES = REPO / "tensor-net_amd" / "tnmps" / "tuning_mi355x.json"


def tuning_for(args, n: int, chi: int, prec: str, batched: bool, groups: int = 1):
    """(label, rows) of the launch-shape table to use, or ("default", None):
    the tuned entry whose workload has this N, chi, precision, batching and
    number of concurrent groups."""
    if args.tune == "none":
        return "default", None
    if args.tune != "auto":
        path, key = args.tune.rsplit(":", 1)
        from tnmps.autotune import load_table
        return key, load_table(path, key)
    if not TUNED_TABLES.exists():
        return "default", None
    data = json.loads(TUNED_TABLES.read_text())
    for key, ent in data.items():
        w = ent["workload"]
        if ((w["n"], w["chi"], w["prec"]) == (n, chi, prec) and (w["traj"] > 1) == batched
                and w.get("groups", 1) == min(groups, w["traj"])):
            return key, ent["table"]
    return "default", None


def contraction_kernel(perf: dict, prec: str, streams: int) -> dict:
    """Achieved matrix-core rate of the contraction kernel k_theta:
    algorithmic flops of the contracted updates (8 m n K, + 32 m n with a
    gate) over the kernel's summed HIP-event spans.  The spans also hold the
    factored updates' elementwise fills, so the rate is a lower bound.  The
    peak is the dense MFMA peak of the arithmetic: FP64 78.6, FP32 157.3 and
    BF16 2500 TFLOP/s (the BF16 split issues three products per real one)."""
    peak = {"f64": FP64_PEAK_TFLOPS, "f32": 157.3, "bf16x3": 2500.0}[prec]
    ms, fl = perf["ms_contract"], perf["contract_flops"]
    tf = fl / (ms / 1e3) / 1e12 if ms > 0 else 0.0
    return {"kernel": "k_theta", "arith": prec, "flops": fl, "ms": round(ms, 3),
            "achieved_TFLOPs": round(tf, 3), "peak_TFLOPs": peak,
            "frac": round(tf / peak, 4), "streams": streams}


def memory_kernels(perf: dict, streams: int) -> dict:
    """Achieved HBM rate of the memory-bound kernels: algorithmic bytes (each
    Gamma element read once and written once, counted by the library from
    every update's m, n and rank and every measured site's chiL, chiR) over
    the kernels' summed HIP-event spans.  With several streams the spans
    overlap other streams' work, so the rates are lower bounds."""
    out = {"peak_GBps": HBM_PEAK_GBS, "streams": streams}
    for name, ms_key, b_key in (("k_split", "ms_split", "bytes_split"),
                                ("k_renorm", "ms_renorm", "bytes_renorm"),
                                ("k_measure", "ms_measure", "bytes_measure")):
        ms, b = perf[ms_key], perf[b_key]
        gbs = b / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        out[name] = {"bytes": b, "ms": round(ms, 3), "GBps": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["k_measure"]["sites"] = perf["measure_sites"]
    return out


def run_mixed(args) -> dict:
    """FP32 contraction vs FP64 on the same outcome record: per-layer error of
    the half-chain entropy and of the Schmidt values, and both update rates."""
    import numpy as np

    from tnmps import Engine
    from tnmps.workload import haar_circuit_fast

    n, chi, T = args.n or 48, args.chi or 512, args.traj or 4
    depth = args.depth or (args.warmup + args.steps)
    circs = [haar_circuit_fast(n, depth, 0.15, args.seed + 11 * i) for i in range(T)]
    out = {}
    records = None
    for prec in ("f64", "f32", "bf16x3"):
        eng = Engine(chi)
        states = [eng.new_qubits(n, ("const", chi)) for _ in range(T)]
        for s in states:
            s.set_precision(prec)
        eng.set_timing(True)
        eng.reset_perf()
        u0, _ = eng.stats()
        t0 = time.perf_counter()
        ents, recs, lams = [], [[] for _ in range(T)], []
        draws = [iter(c.draws) for c in circs]
        for t in range(depth):
            if records is None:
                res = eng.batch_bilayer(states, [c.layers[t] for c in circs], draws=draws)
            else:
                res = eng.batch_bilayer(states, [c.layers[t] for c in circs],
                                        forced=[records[i][t] for i in range(T)])
            for i, r in enumerate(res):
                recs[i].append(r[0])
            ents.append([r[2] for r in res])
            lams.append([s.schmidt(n // 2 - 1) for s in states])
            log(f"{prec} layer {t + 1}/{depth} ({time.perf_counter() - t0:.1f}s), "
                f"max bond {max(s.max_bond_dim() for s in states)}")
        eng.sync()
        wall = time.perf_counter() - t0
        u1, _ = eng.stats()
        out[prec] = {"ents": np.array(ents), "lams": lams, "rate": (u1 - u0) / wall,
                     "wall": wall, "perf": eng.perf()}
        if records is None:
            records = recs
        del states, eng
    def errors(prec):
        ent_err = np.max(np.abs(out[prec]["ents"] - out["f64"]["ents"]), axis=1)
        lam_err = []
        for t in range(depth):
            m = 0.0
            for a, b in zip(out[prec]["lams"][t], out["f64"]["lams"][t]):
                k = min(len(a), len(b))
                m = max(m, float(np.max(np.abs(a[:k] - b[:k]))) if k else 0.0)
            lam_err.append(m)
        return ent_err, lam_err

    ent_err, lam_err = errors("f32")
    bf_ent, bf_lam = errors("bf16x3")
    return {
        "metric": "two-site updates/s, FP32 contraction vs FP64 (mixed precision)",
        "value": round(out["f32"]["rate"], 3), "unit": "two-site updates/s", "n_gpus": 1,
        "steps": depth, "warmup": 0, "ms_per_step": round(out["f32"]["wall"] / depth * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 contraction, f64 SVD", "data": "synthetic",
        "config": {"workload": f"{T} trajectories N={n} depth={depth} p=0.15 Const({chi}); the FP32 "
                               f"run replays the FP64 run's outcome record",
                   "fp64_updates_per_s": round(out["f64"]["rate"], 3),
                   "entropy_abs_err_per_layer": [float(f"{x:.3e}") for x in ent_err],
                   "schmidt_abs_err_per_layer": [float(f"{x:.3e}") for x in lam_err],
                   "max_entropy_err": float(f"{float(np.max(ent_err)):.3e}"),
                   "bf16x3_updates_per_s": round(out["bf16x3"]["rate"], 3),
                   "bf16x3_entropy_abs_err_per_layer": [float(f"{x:.3e}") for x in bf_ent],
                   "bf16x3_schmidt_abs_err_per_layer": [float(f"{x:.3e}") for x in bf_lam],
                   "bf16x3_max_entropy_err": float(f"{float(np.max(bf_ent)):.3e}"),
                   "contraction_kernel": {p: contraction_kernel(out[p]["perf"], p, 1)
                                          for p in out}},
        "roofline": None,
    }


def run_gates(args) -> dict:
    """Gate layers only, on saturated trajectories of one engine: every update
    is a full two-site gate update (theta of 2chi x 2chi, C = 2chi), so the
    contraction kernel's event span holds contractions only.  Reports the
    update rate and the contraction kernel's matrix-core rate."""
    from tnmps import Engine
    from tnmps.workload import haar_circuit_fast

    n, chi, T = args.n or 64, args.chi or 256, args.traj or 64
    W, K = args.warmup, args.steps
    circs = [haar_circuit_fast(n, W + K, args.p, args.seed + 7 * i) for i in range(T)]
    eng = Engine(chi)
    states = [eng.new_qubits(n, ("const", chi)) for _ in range(T)]
    for st in states:
        st.set_precision(args.prec)
    draws = [iter(c.draws) for c in circs]
    t_w = time.perf_counter()
    for t in range(W):
        eng.batch_bilayer(states, [c.layers[t] for c in circs], draws=draws)
        log(f"warmup layer {t + 1}/{W} ({time.perf_counter() - t_w:.1f}s)")
    eng.sync()
    eng.set_timing(True)
    eng.reset_perf()
    u0, _ = eng.stats()
    t0 = time.perf_counter()
    for t in range(W, W + K):
        eng.batch_bilayer(states, [(c.layers[t][0], []) for c in circs], forced=[[]] * T)
    eng.sync()
    wall = time.perf_counter() - t0
    eng.set_timing(False)
    u1, _ = eng.stats()
    perf = eng.perf()
    return {
        "metric": "two-site gate updates/s (gate layers only)",
        "value": round((u1 - u0) / wall, 2), "unit": "two-site updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(wall / K * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.prec == "f64" else f"{args.prec} contraction, f64 SVD",
        "data": "synthetic",
        "config": {"workload": f"{T} saturated trajectories N={n} Const({chi}), brickwork gate "
                               f"layers without measurements, one engine",
                   "trajectories": T, "n_qubits": n, "chi_max": chi,
                   "updates_timed": u1 - u0,
                   "max_bond": max(s.max_bond_dim() for s in states)},
        "roofline": None,
        "breakdown": {k: round(perf[k], 2) for k in ("ms_theta", "ms_jacobi", "ms_post",
                                                     "ms_contract")},
        "contraction_kernel": contraction_kernel(perf, args.prec, 1),
    }


def run_serving(args) -> dict:
    """Poisson stream of trajectory requests served by continuous batching."""
    import numpy as np

    from tnmps.serve import Request, Server, percentile
    from tnmps.workload import haar_circuit_fast

    rng = np.random.default_rng(args.seed)
    ns = [int(x) for x in args.serve_n.split(",")]
    chis = [int(x) for x in args.serve_chi.split(",")]
    rate, dur = args.rate, args.duration
    reqs, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > dur:
            break
        n = int(rng.choice(ns))
        p = float(rng.uniform(0.0, 0.4))
        chi = int(rng.choice(chis))
        if args.serve_host_circuits:
            c = haar_circuit_fast(n, n, p, int(rng.integers(1 << 30)))
            reqs.append(Request(rid=len(reqs), n=n, p=p, chi=chi, layers=c.layers,
                                draws=c.draws, arrival=t, prec="f32"))
        else:
            # gates, measurement sites and draws generated on the device
            reqs.append(Request(rid=len(reqs), n=n, p=p, chi=chi, arrival=t, prec="f32",
                                gen_seed=args.seed, traj_id=len(reqs), gen_depth=n))
    log(f"{len(reqs)} requests over {dur}s (rate {rate}/s)")
    srv = Server(chi_max=max(chis), hbm_budget=args.hbm_budget * 1e9, log=log,
                 workers=args.serve_workers, dispatch=args.serve_dispatch)
    u0, _ = srv.stats()
    t0 = time.perf_counter()
    srv.run(reqs)
    wall = time.perf_counter() - t0
    u1, _ = srv.stats()
    lat = [r.latency for r in reqs if r.latency is not None]
    return {
        "metric": "trajectories/s + p50/p99 latency (continuous batching)",
        "value": round(len(lat) / wall, 4), "unit": "trajectories/s", "n_gpus": 1,
        "steps": srv.ticks, "warmup": 0, "ms_per_step": round(wall / max(srv.ticks, 1) * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 contraction, f64 SVD", "data": "synthetic",
        "config": {"workload": f"Poisson arrivals {rate}/s for {dur}s, N in {ns}, p~U[0,0.4], "
                               f"chi_max in {chis}, depth=N, Haar brickwork + Z measurements, "
                               f"{args.serve_workers} lock-step worker(s), "
                               f"{args.serve_dispatch} dispatch",
                   "circuits": "host-generated" if args.serve_host_circuits
                               else "generated on the device (Philox4x64-10 streams)",
                   "requests": len(reqs), "completed": len(lat),
                   "p50_latency_s": round(percentile(lat, 50), 3),
                   "p99_latency_s": round(percentile(lat, 99), 3),
                   "mean_latency_s": round(float(np.mean(lat)), 3) if lat else None,
                   "updates_per_s": round((u1 - u0) / wall, 2),
                   "trajectory_layers_per_s": round(srv.layers_done / wall, 2),
                   "peak_resident_bytes": srv.peak_held, "wall_s": round(wall, 2)},
        "roofline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--chi", type=int, default=0)
    ap.add_argument("--p", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=10546)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="CPU-baseline processes (0: one per core, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="single: leave out the batched FP32 and FP64 measurement (mixed_batched)")
    ap.add_argument("--mixed-traj", type=int, default=256,
                    help="single: trajectories of the batched measurement")
    ap.add_argument("--mixed-steps", type=int, default=1,
                    help="single: timed layers of each batched measurement")
    ap.add_argument("--workload", default="single",
                    choices=["single", "batched", "mixed", "serving", "gates"])
    ap.add_argument("--traj", type=int, default=0, help="trajectories (batched, mixed)")
    ap.add_argument("--depth", type=int, default=0, help="circuit depth (batched, mixed)")
    ap.add_argument("--pgrid", default="0.05,0.30,11",
                    help="batched: measurement rates linspace(lo,hi,count), dealt round-robin")
    ap.add_argument("--prec", default="f64", choices=["f64", "f32", "bf16x3"],
                    help="batched: contraction precision (the SVD is FP64 either way)")
    ap.add_argument("--groups", type=int, default=4,
                    help="batched: lock-step groups, one engine, stream and host thread each "
                         "(4 = the hardware queues HIP gives a process)")
    ap.add_argument("--rate", type=float, default=2.0, help="serving: mean arrivals per s")
    ap.add_argument("--duration", type=float, default=20.0, help="serving: arrival window s")
    ap.add_argument("--serve-n", default="16,32,64,128")
    ap.add_argument("--serve-chi", default="64,128,256")
    ap.add_argument("--hbm-budget", type=float, default=160.0,
                    help="serving: GB of device memory for MPS state and the workers' workspaces")
    ap.add_argument("--serve-workers", type=int, default=4,
                    help="serving: lock-step loops, one engine, stream and host thread each")
    ap.add_argument("--serve-dispatch", default="least_loaded",
                    choices=["least_loaded", "round_robin"],
                    help="serving: how requests are dealt to the workers")
    ap.add_argument("--serve-host-circuits", action="store_true",
                    help="serving: requests carry host-generated circuits")
    ap.add_argument("--tune", default="auto",
                    help="launch-shape table: 'auto' (the package's tuned table for this "
                         "workload's N, chi and precision, if any), 'none', or PATH:KEY "
                         "(a file written by python -m tnmps.autotune)")
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

    if args.workload != "single":
        fn = {"batched": run_batched, "mixed": run_mixed, "serving": run_serving,
              "gates": run_gates}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return 0

    # the CPU baseline's workers fork before this process touches the GPU
    pool, workers = (None, 0) if args.no_cpu else start_cpu_pool(args.cpu_workers)
    host = cpu_info()
    try:
        out = run_single(args, pool, workers, host)
    finally:
        if pool is not None:
            pool.close()
            pool.join()
    print(json.dumps(out), flush=True)
    return 0


def run_single(args, pool, workers: int, host: dict) -> dict:
    """The metric's line: the FP64 single trajectory, then the batched path
    of the same N, chi and p in FP32 and FP64 (mixed_batched), then the CPU
    baseline."""
    from tnmps import MPS
    from tnmps.workload import haar_circuit

    n, chi, p = args.n or 64, args.chi or 256, args.p
    W, K = args.warmup, args.steps
    W_cap = max(W, 3 * n)  # saturation comes long before this
    trunc = ("const", chi)
    c = haar_circuit(n, W_cap + K + 1, p, args.seed)
    draws = iter(c.draws)
    st = MPS.new_qubits(n, trunc)
    tune_label, tune_rows = tuning_for(args, n, chi, "f64", batched=False)
    if tune_rows is not None:
        st.set_tuning(tune_rows)
    t_w = time.perf_counter()
    W_run = 0
    while W_run < W_cap and (W_run < W or not saturated([st], chi)):
        st.run_circuit(c.layers[W_run:W_run + 1], rs=draws)
        W_run += 1
        log(f"warmup layer {W_run}: max bond {st.max_bond_dim()} "
            f"({time.perf_counter() - t_w:.1f}s)")
    sat = saturated([st], chi)
    st.sync()
    st.set_timing(True)
    st.reset_perf()
    u0, _ = st.stats()
    t0 = time.perf_counter()
    for t in range(W_run, W_run + K):
        st.run_circuit(c.layers[t:t + 1], rs=draws)
        log(f"timed layer {t - W_run + 1}/{K}: max bond {st.max_bond_dim()} "
            f"({time.perf_counter() - t0:.1f}s)")
    st.sync()
    t1 = time.perf_counter()
    st.set_timing(False)
    u1, _ = st.stats()
    perf = st.perf()
    updates = u1 - u0
    wall = t1 - t0
    value = updates / wall

    # The dominant kernels, the SVD's (the QR preconditioner's kernels and the
    # Jacobi step kernels: one-launch row-split steps, or one workgroup per
    # block pair), are credited with the algorithmic R-SVD flops of the SVDs
    # they perform, over the HIP-event spans of the QR and of the Jacobi
    # sweeps on the library's stream (dispatch gaps between dependent
    # launches included).  The traffic is that of the most frequent step
    # dispatch, a lone refactor's row-split step.
    svd_s = (perf["ms_jacobi"] + perf["ms_qr"]) / 1e3
    achieved = perf["svd_flops"] / svd_s / 1e12 if svd_s > 0 else 0.0
    roof = {
        "bound": "mfma",
        "kernel": "SVD: QR preconditioner (k_qr_*) + Jacobi steps (k_jacobi_fused_r<4, 1> "
                  "row-split, k_jacobi_step<8> gate batches)",
        "achieved": round(achieved, 4),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
        "traffic": round(TRAFFIC_BYTES["k_jacobi_fused_r<4, 1>"]),
        "traffic_source": TRAFFIC_SOURCE,
        "jacobi_only_TFLOPs": round(perf["svd_flops"] / (perf["ms_jacobi"] / 1e3) / 1e12, 4)
        if perf["ms_jacobi"] > 0 else None,
        "avg_launch_us": round(perf["ms_jacobi"] * 1e3 / max(perf["jacobi_launches"], 1), 3),
        "launches": perf["jacobi_launches"],
        "svd_flops": perf["svd_flops"],
    }
    breakdown = {k: round(perf[k], 3) for k in ("ms_theta", "ms_qr", "ms_jacobi", "ms_post",
                                                "ms_measure", "ms_host_launch", "qr_tasks")}
    breakdown["sweeps_per_update_batch"] = round(perf["sweeps"] / max(perf["updates"], 1), 3)
    breakdown["alg_tflops_whole_update"] = round(perf["alg_flops"] / wall / 1e12, 4)

    mixed = None
    if not args.no_mixed:
        mixed = mixed_batched(args, value)

    cpu = None
    if pool is not None:
        rate, cnt, el = cpu_baseline(pool, workers, st, n, trunc, c.layers[W_run + K],
                                     args.cpu_budget)
        cpu = {"value": round(rate, 4), "unit": "two-site updates/s", "cores": workers,
               "kind": "port",
               "sample": f"oracle (numpy + LAPACK zgesvd, 1 thread per process) in {workers} "
                         f"processes, one per core, each on its own copy of the saturated "
                         f"N={n} chi={chi} state: the next layer's gates, then refactors, "
                         f"for {args.cpu_budget:.0f}s; {cnt} updates in all",
               "host": host}
        log(f"cpu baseline: {rate:.3f} updates/s over {cnt} updates on {workers} cores")

    return {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "two-site updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W_run,
        "ms_per_step": round(wall / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"single trajectory N={n} Haar brickwork, p={p} Z measurements, "
                        f"Const(chi_max={chi}), complex128, plain FP64 GPU path; "
                        + ("step = one saturated circuit layer" if sat else
                           "step = one circuit layer (the state did not saturate)"),
            "n_qubits": n, "p": p, "chi_max": chi, "trajectories": 1,
            "updates_timed": updates, "tuning": tune_label, "saturated": sat,
            "warmup_layers_min": W,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed_batched": mixed,
        "breakdown": breakdown,
        "memory_kernels": memory_kernels(perf, 1),
        "contraction_kernel": contraction_kernel(perf, "f64", 1),
    }


def mixed_batched(args, single_value: float) -> dict:
    """The batched path at the single line's N, chi and p: --mixed-traj
    trajectories in 4 lock-step groups, first with the FP32 contraction
    (warmed up as the single line is, until saturated), then with FP64
    throughout on the same circuits and the same number of warmup layers, so
    both time the same layers; --mixed-steps timed layers each.  Each rate's
    ratio to this run's FP64 single-trajectory value, and the FP32 run's
    half-chain entropy error against an FP64 replay of its first
    trajectories' outcome records."""
    import copy

    a = copy.copy(args)
    a.n, a.chi = args.n or 64, args.chi or 256
    a.traj, a.groups = args.mixed_traj, 4
    a.pgrid = f"{args.p},{args.p},1"
    a.depth = a.n
    a.steps = args.mixed_steps
    legs = {}
    warm = 1
    for prec in ("f32", "f64"):
        a.prec, a.warmup = prec, warm
        log(f"batched leg: {a.traj} trajectories N={a.n} chi={a.chi} p={args.p} {prec}")
        r = run_batched(a)
        warm = r["warmup"]
        leg = {"value": r["value"], "unit": "two-site updates/s",
               "ratio_vs_fp64_single": round(r["value"] / single_value, 3) if single_value > 0
               else None,
               "jacobi_achieved_TFLOPs": r["roofline"]["achieved"],
               "saturated": r["config"]["saturated"]}
        if prec == "f32":
            leg["fp64_replay"] = r["config"]["fp64_replay"]
            leg["memory_kernels"] = r["memory_kernels"]
        legs[prec] = leg
    rep = legs["f32"]["fp64_replay"]
    err = rep["max_half_chain_entropy_abs_err_vs_fp64"] if rep else None
    ratio = legs["f32"]["ratio_vs_fp64_single"]
    return {
        "workload": f"{a.traj} trajectories N={a.n} depth={a.depth} p={args.p} Const({a.chi}), "
                    f"4 lock-step groups on 4 streams; step = one layer of every trajectory "
                    f"after {warm} warmup layers (saturated)",
        "trajectories": a.traj, "groups": 4, "warmup": warm, "steps": a.steps,
        "f32": legs["f32"], "f64": legs["f64"],
        "value": legs["f32"]["value"], "ratio_vs_fp64_single": ratio,
        "target": ">= 20x the FP64 single-trajectory value with entropy error <= 1e-6",
        "target_met": bool(ratio and ratio >= 20.0 and err is not None and err <= 1e-6),
    }


if __name__ == "__main__":
    sys.exit(main())
