"""Launch-shape autotuner for the update path.

An engine's tuning table (``tn_tune`` in include/tnmps.h) picks, per bucket
of the SVD column count C, the contraction tile, the Jacobi block width, the
row split of lone SVDs and how many sweeps the host enqueues ahead.  Every
choice computes the same SVD to rounding; only the time changes.  No
counterpart in the reference, which makes one LAPACK call per update.

The tuner is driven by a workload, not by synthetic kernels: it brings a set
of trajectories (N qubits, Haar brickwork, measurement rate p, Const(chi))
to saturation, records the outcomes of a few further layers, and then times
those layers from clones of the saturated states, replayed on the recorded
outcomes, once per candidate table.  Every candidate therefore does the same
updates on the same inputs.  It tunes the buckets that hold the saturated
updates, 2 chi (gates and full refactors, theta of 2chi x 2chi) and chi
(factored refactors, one factor of 2chi x chi), by coordinate descent: one
field at a time, keeping a value only when it is faster by more than
``min_gain``.

Command line (runs on the device; writes the table as JSON):

    python -m tnmps.autotune --n 64 --chi 256 --p 0.1 --traj 1 --out table.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Dict, List, Sequence

from . import _lib as L
from .mps import Engine
from .workload import haar_circuit_fast

# candidate values per field, in the order the descent visits them
CANDIDATES: Dict[str, Sequence[int]] = {
    "theta_tile": (32, 64),
    "split_max": (1, 2, 4, 8),
    "split_rows": (64, 128, 256),
    "bw": (8, 16),
    "ahead": (1, 2, 3),
}


def bucket_of(table: List[dict], C: int) -> int:
    """Index of the row that serves an update batch of largest column count C."""
    for i, r in enumerate(table):
        if r["c_max"] >= C:
            return i
    return len(table) - 1


class _Workload:
    """Saturated trajectories on one engine plus a recorded stretch of layers."""

    def __init__(self, n, chi, p, traj, warm, layers, seed, prec, log):
        self.eng = Engine(chi)
        self.n, self.chi = n, chi
        depth = warm + layers
        circs = [haar_circuit_fast(n, depth, p, seed + 11 * i) for i in range(traj)]
        self.states = [self.eng.new_qubits(n, ("const", chi)) for _ in range(traj)]
        for st in self.states:
            st.set_precision(prec)
        draws = [iter(c.draws) for c in circs]
        t0 = time.perf_counter()
        for t in range(warm):
            self.eng.batch_bilayer(self.states, [c.layers[t] for c in circs], draws=draws)
        self.eng.sync()
        log(f"autotune: {traj} x N={n} chi={chi} p={p} saturated in "
            f"{time.perf_counter() - t0:.1f}s, max bond "
            f"{max(s.max_bond_dim() for s in self.states)}")
        self.layers = [[c.layers[t] for c in circs] for t in range(warm, depth)]
        # the outcomes of the timed stretch, from one sampled pass on clones
        clones = [s.clone() for s in self.states]
        self.forced = [[] for _ in range(traj)]
        for lay in self.layers:
            res = self.eng.batch_bilayer(clones, lay, draws=draws)
            for i, r in enumerate(res):
                self.forced[i].append(r[0])
        self.eng.sync()
        del clones

    def time_table(self, table: List[dict], repeat: int) -> float:
        """Best wall time of the recorded stretch under ``table``."""
        self.eng.set_tuning(table)
        best = float("inf")
        for _ in range(repeat):
            clones = [s.clone() for s in self.states]
            self.eng.sync()
            t0 = time.perf_counter()
            for t, lay in enumerate(self.layers):
                self.eng.batch_bilayer(clones, lay, forced=[f[t] for f in self.forced])
            self.eng.sync()
            best = min(best, time.perf_counter() - t0)
            del clones
        return best


def tune(n: int = 64, chi: int = 256, p: float = 0.1, traj: int = 1, warm: int = 0,
         layers: int = 2, repeat: int = 2, seed: int = 10546, prec: str = "f64",
         min_gain: float = 0.02, log=print) -> dict:
    """Tune the table rows of buckets 2chi and chi on the workload described.

    ``warm`` layers bring the chain to saturation (0: 2 log2(chi) + 4).
    Returns {"table": rows, "trials": [...], "workload": {...}} where trials
    holds every timed candidate (bucket, field, value, seconds)."""
    if warm <= 0:
        warm = 2 * max(1, int(chi).bit_length()) + 4
    w = _Workload(n, chi, p, traj, warm, layers, seed, prec, log)
    table = w.eng.get_tuning()
    base = w.time_table(table, repeat)
    trials = [{"bucket": None, "field": None, "value": None, "s": round(base, 5)}]
    log(f"autotune: default table {base * 1e3:.1f} ms")
    best = base
    for C in (2 * chi, chi):
        bi = bucket_of(table, C)
        if any(t["bucket"] == table[bi]["c_max"] for t in trials[1:]):
            continue  # both sizes fall into one row
        for field, values in CANDIDATES.items():
            for v in values:
                if table[bi][field] == v:
                    continue
                cand = [dict(r) for r in table]
                cand[bi][field] = v
                s = w.time_table(cand, repeat)
                trials.append({"bucket": table[bi]["c_max"], "field": field, "value": v,
                               "s": round(s, 5)})
                log(f"autotune: c_max {table[bi]['c_max']} {field}={v}: {s * 1e3:.1f} ms "
                    f"(best {best * 1e3:.1f})")
                if s < best * (1.0 - min_gain):
                    table, best = cand, s
    w.eng.set_tuning(table)
    return {"table": table, "default_s": round(base, 5), "tuned_s": round(best, 5),
            "trials": trials,
            "workload": {"n": n, "chi": chi, "p": p, "traj": traj, "warm": warm,
                         "layers": layers, "repeat": repeat, "seed": seed, "prec": prec}}


def load_table(path: str, key: str) -> List[dict]:
    """Rows of entry ``key`` of a JSON file written by :func:`main`."""
    with open(path) as f:
        data = json.load(f)
    rows = data[key]["table"]
    for r in rows:
        if set(r) != set(L.TUNE_FIELDS):
            raise ValueError(f"{path}:{key}: row fields {sorted(r)}")
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--chi", type=int, default=256)
    ap.add_argument("--p", type=float, default=0.1)
    ap.add_argument("--traj", type=int, default=1)
    ap.add_argument("--warm", type=int, default=0)
    ap.add_argument("--layers", type=int, default=2)
    ap.add_argument("--repeat", type=int, default=2)
    ap.add_argument("--prec", default="f64", choices=["f64", "f32", "bf16x3"])
    ap.add_argument("--key", default=None, help="entry name in the output file")
    ap.add_argument("--out", required=True, help="JSON file (entries are merged)")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    res = tune(a.n, a.chi, a.p, a.traj, a.warm, a.layers, a.repeat, prec=a.prec, log=log)
    key = a.key or f"traj{a.traj}_n{a.n}_chi{a.chi}_{a.prec}"
    try:
        with open(a.out) as f:
            data = json.load(f)
    except FileNotFoundError:
        data = {}
    data[key] = res
    with open(a.out, "w") as f:
        json.dump(data, f, indent=1)
    log(f"autotune: {key}: {res['default_s'] * 1e3:.1f} -> {res['tuned_s'] * 1e3:.1f} ms")
    return 0


if __name__ == "__main__":
    sys.exit(main())
