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
import threading
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
    "qr": (0, 1),
}


def bucket_of(table: List[dict], C: int) -> int:
    """Index of the row that serves an update batch of largest column count C."""
    for i, r in enumerate(table):
        if r["c_max"] >= C:
            return i
    return len(table) - 1


class _Workload:
    """Saturated trajectories plus a recorded stretch of layers, dealt
    round-robin into ``groups`` lock-step groups, one engine each; the groups
    run in their own host threads at once, as the batched driver runs them."""

    def __init__(self, n, chi, p, traj, warm, layers, seed, prec, log, groups=1):
        G = max(1, min(groups, traj))
        self.engs = [Engine(chi) for _ in range(G)]
        self.n, self.chi = n, chi
        depth = warm + layers
        circs = [haar_circuit_fast(n, depth, p, seed + 11 * i) for i in range(traj)]
        self.idx = [list(range(g, traj, G)) for g in range(G)]
        self.states = [[self.engs[g].new_qubits(n, ("const", chi)) for _ in ix]
                       for g, ix in enumerate(self.idx)]
        for grp in self.states:
            for st in grp:
                st.set_precision(prec)
        draws = [iter(c.draws) for c in circs]
        t0 = time.perf_counter()
        for g, ix in enumerate(self.idx):
            for t in range(warm):
                self.engs[g].batch_bilayer(self.states[g], [circs[i].layers[t] for i in ix],
                                           draws=[draws[i] for i in ix])
            self.engs[g].sync()
        log(f"autotune: {traj} x N={n} chi={chi} p={p} in {G} group(s) saturated in "
            f"{time.perf_counter() - t0:.1f}s, max bond "
            f"{max(s.max_bond_dim() for grp in self.states for s in grp)}")
        self.layers = [[[circs[i].layers[t] for i in ix] for t in range(warm, depth)]
                       for ix in self.idx]
        # the outcomes of the timed stretch, from one sampled pass on clones
        self.forced = []
        for g, ix in enumerate(self.idx):
            clones = [s.clone() for s in self.states[g]]
            rec = [[] for _ in ix]
            for lay in self.layers[g]:
                res = self.engs[g].batch_bilayer(clones, lay, draws=[draws[i] for i in ix])
                for j, r in enumerate(res):
                    rec[j].append(r[0])
            self.engs[g].sync()
            self.forced.append(rec)
            del clones

    def time_table(self, table: List[dict], repeat: int) -> float:
        """Best wall time of the recorded stretch under ``table`` (every group
        at once)."""
        for e in self.engs:
            e.set_tuning(table)
        best = float("inf")
        for _ in range(repeat):
            clones = [[s.clone() for s in grp] for grp in self.states]
            for e in self.engs:
                e.sync()
            errors: list = []

            def run(g):
                try:
                    for t, lay in enumerate(self.layers[g]):
                        self.engs[g].batch_bilayer(clones[g], lay,
                                                   forced=[f[t] for f in self.forced[g]])
                    self.engs[g].sync()
                except Exception as exc:  # surfaced after the join
                    errors.append(exc)

            t0 = time.perf_counter()
            th = [threading.Thread(target=run, args=(g,)) for g in range(len(self.engs))]
            for x in th:
                x.start()
            for x in th:
                x.join()
            if errors:
                raise errors[0]
            best = min(best, time.perf_counter() - t0)
            del clones
        return best


def tune(n: int = 64, chi: int = 256, p: float = 0.1, traj: int = 1, warm: int = 0,
         layers: int = 2, repeat: int = 2, seed: int = 10546, prec: str = "f64",
         min_gain: float = 0.02, log=print, groups: int = 1) -> dict:
    """Tune the table rows of buckets 2chi and chi on the workload described.

    ``warm`` layers bring the chain to saturation (0: 2 log2(chi) + 4).
    Returns {"table": rows, "trials": [...], "workload": {...}} where trials
    holds every timed candidate (bucket, field, value, seconds)."""
    if warm <= 0:
        warm = 2 * max(1, int(chi).bit_length()) + 4
    w = _Workload(n, chi, p, traj, warm, layers, seed, prec, log, groups)
    table = w.engs[0].get_tuning()
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
    for e in w.engs:
        e.set_tuning(table)
    return {"table": table, "default_s": round(base, 5), "tuned_s": round(best, 5),
            "trials": trials,
            "workload": {"n": n, "chi": chi, "p": p, "traj": traj, "warm": warm,
                         "layers": layers, "repeat": repeat, "seed": seed, "prec": prec,
                         "groups": len(w.engs)}}


def load_table(path: str, key: str) -> List[dict]:
    """Rows of entry ``key`` of a JSON file written by :func:`main`."""
    with open(path) as f:
        data = json.load(f)
    rows = data[key]["table"]
    out = []
    for r in rows:
        missing = set(L.TUNE_FIELDS) - set(r)
        if set(r) - set(L.TUNE_FIELDS) or missing - set(L.TUNE_DEFAULTS):
            raise ValueError(f"{path}:{key}: row fields {sorted(r)}")
        # a table written before a field existed takes that field's default
        out.append({f: r[f] if f in r else L.TUNE_DEFAULTS[f] for f in L.TUNE_FIELDS})
    return out


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
    ap.add_argument("--groups", type=int, default=1,
                    help="lock-step groups run at once, one engine and host thread each")
    ap.add_argument("--key", default=None, help="entry name in the output file")
    ap.add_argument("--out", required=True, help="JSON file (entries are merged)")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    res = tune(a.n, a.chi, a.p, a.traj, a.warm, a.layers, a.repeat, prec=a.prec, log=log,
               groups=a.groups)
    key = a.key or f"traj{a.traj}_n{a.n}_chi{a.chi}_{a.prec}_g{res['workload']['groups']}"
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
