"""Co-evolution drivers: untruncated trajectories whose sampled outcomes are
replayed on truncated copies, recording each outcome's probability at every
bond dimension (workspace/src/haar_coev_probs.rs,
workspace/src/haar_coev_probs_batching.rs).

The reference runs one state per rayon thread.  Here every run and every
replica of a circuit is a handle on one engine: the bilayer driver advances
all untruncated runs in lock step, then all (run, bond dimension) replicas
in lock step, so one launch carries a whole batch of two-site updates.

Uniform draws for sampled measurements come from a numpy generator (the
reference uses ``thread_rng``), consumed as the reference consumes its
generator: one draw per sampled measurement, in processing order.
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import formats as F
from .mps import Engine

# haar_coev_probs.rs:17-23
BONDS_COEV: Tuple[Optional[int], ...] = (4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 44, 48, 56, 64,
                                         72, None)
# haar_coev_probs_batching.rs:26-30
BONDS_BATCHING: Tuple[Optional[int], ...] = (None, 4, 6, 8, 10, 12, 16, 20, 24, 28, 32)


def exact_bond_cap(n: int) -> int:
    """Largest untruncated bond dimension of n qubits, 2^floor(n/2)."""
    return 1 << (n // 2)


def _uni_mats(layer: Sequence[F.Uni], rng) -> List[Tuple[int, np.ndarray]]:
    """A unitary layer as (k, matrix) pairs (``apply_uni_rng``,
    mps.rs:2110-2126)."""
    out = []
    for u in layer:
        if isinstance(u, F.Mat):
            out.append((int(u.k), np.asarray(u.mat, dtype=np.complex128)))
        else:
            out.append((int(u.k), u.matrix(rng)))
    return out


def _meas_ops(layer: Sequence[F.Meas]) -> list:
    return [m.k if isinstance(m, F.Rand) else (m.k, m.outcome) for m in layer]


def coev_probs(nqubits: int, unis: Sequence[Sequence[F.Uni]], meas: Sequence[Sequence[F.Meas]],
               runs: int = 64, bonds: Sequence[Optional[int]] = BONDS_COEV,
               rng: Optional[np.random.Generator] = None,
               exact_cap: Optional[int] = None) -> Dict[str, np.ndarray]:
    """``sample_trajectory`` and the record assembly of haar_coev_probs.rs
    (:33-83, :163-203) for one circuit and measurement layer list.

    Per run: an untruncated state (``BondDim`` None, bond dimensions up to
    ``exact_cap``) runs ``apply_bilayer`` on every layer with sampled
    measurements; then a ``Const(chi)`` state per chi in ``bonds`` replays
    the outcomes as projections.  Returns ``meas_locs`` (num_meas x 2 u8:
    layer, qubit, from the layer list), ``traj_data`` (runs x num_meas u8
    outcomes) and ``prob_data`` (runs x len(bonds) x num_meas f64: row b the
    probabilities under ``bonds[b]``, the untruncated state's for None).

    ``Uni::Gate`` unitaries are turned into matrices once per circuit, so a
    ``Haar2`` gate is the same matrix on every run and replica (the reference
    draws it anew on each evaluation)."""
    if len(unis) != len(meas):
        raise ValueError("unitary and measurement layer counts differ")
    rng = rng if rng is not None else np.random.default_rng()
    n = int(nqubits)
    depth = len(unis)
    meas_locs = np.asarray([(t, m.k) for t, layer in enumerate(meas) for m in layer],
                           dtype=np.uint8).reshape(-1, 2)
    num_meas = meas_locs.shape[0]
    mats = [_uni_mats(layer, rng) for layer in unis]
    mops = [_meas_ops(layer) for layer in meas]
    chis = [c for c in bonds if c is not None]
    cap = int(exact_cap) if exact_cap is not None else exact_bond_cap(n)
    eng = Engine(max([cap] + chis))

    exact = [eng.new_qubits(n, None, chi_cap=cap) for _ in range(runs)]
    draws = [iter(rng.random(num_meas + n * depth)) for _ in range(runs)]
    rec: List[List[List[Tuple[int, int]]]] = [[] for _ in range(runs)]
    pq: List[List[float]] = [[] for _ in range(runs)]
    for t in range(depth):
        res = eng.batch_bilayer(exact, [(mats[t], mops[t])] * runs, draws=draws)
        for r, (record, probs, _) in enumerate(res):
            rec[r].append(record)
            pq[r] += [p for _, p in probs]
    for r in range(runs):
        if sum(len(layer) for layer in rec[r]) != num_meas:
            raise ValueError("measurement layers must list distinct in-range qubits "
                             "(the outcome record must match meas_locs)")
    del exact

    traj = np.asarray([[o for layer in rec[r] for _, o in layer] for r in range(runs)],
                      dtype=np.uint8).reshape(runs, num_meas)
    prob = np.zeros((runs, len(bonds), num_meas))
    for r in range(runs):
        prob[r, [b for b, c in enumerate(bonds) if c is None], :] = pq[r]
    if chis:
        reps = [(r, b) for r in range(runs) for b, c in enumerate(bonds) if c is not None]
        states = [eng.new_qubits(n, ("const", bonds[b])) for (_, b) in reps]
        got: List[List[float]] = [[] for _ in reps]
        for t in range(depth):
            res = eng.batch_bilayer(states, [(mats[t], [])] * len(reps),
                                    forced=[rec[r][t] for (r, _) in reps])
            for i, (_, probs, _) in enumerate(res):
                got[i] += [p for _, p in probs]
        for i, (r, b) in enumerate(reps):
            prob[r, b, :] = got[i]
    return {"meas_locs": meas_locs, "traj_data": traj, "prob_data": prob}


def coev_probs_fname(man: F.MiptManifest, p: float, circ: int, output_id: str) -> str:
    """haar_coev_probs.rs:205-222."""
    return (f"haar_coev_probs_seed={man.seed}_nqubits={man.nqubits}_depth={man.depth}"
            f"_p={p:.6f}_circ={circ}_outid={output_id}.npz")


def save_coev_probs(outdir, man: F.MiptManifest, manifest_file: str, p: float, circ: int,
                    output_id: str, res: Dict[str, np.ndarray]) -> str:
    """The output file of haar_coev_probs.rs:223-242 (text fields as i32
    character codes)."""
    path = os.path.join(outdir, coev_probs_fname(man, p, circ, output_id))
    codes = lambda s: np.asarray([ord(c) for c in s], dtype=np.int32)
    np.savez(path,
             manifest_file=codes(str(manifest_file)),
             seed=np.asarray([man.seed], dtype=np.uint64),
             nqubits=np.asarray([man.nqubits], dtype=np.int32),
             depth=np.asarray([man.depth], dtype=np.int32),
             num_circs=np.asarray([man.num_circs], dtype=np.int32),
             genid=codes(man.id),
             p=np.asarray([p], dtype=np.float64),
             circ=np.asarray([circ], dtype=np.int32),
             meas_locs=res["meas_locs"], traj_data=res["traj_data"], prob_data=res["prob_data"])
    return path


def run_manifest(manifest_file, circ: int, p: float, runs: int = 64,
                 bonds: Sequence[Optional[int]] = BONDS_COEV,
                 rng: Optional[np.random.Generator] = None,
                 exact_cap: Optional[int] = None) -> Tuple[F.MiptManifest, Dict[str, np.ndarray]]:
    """haar_coev_probs.rs ``main`` (:94-161) up to the output: the circuit
    and layer files sit beside the manifest."""
    man = F.MiptManifest.load(manifest_file)
    if circ >= man.num_circs:
        raise ValueError(f"invalid circuit number {circ} for {man.num_circs} total circuits")
    if p not in man.p_meas:
        raise ValueError(f"invalid p value {p:.6f} for batch values {man.p_meas}")
    d = os.path.dirname(os.path.abspath(manifest_file))
    unis = F.load_uni_circuit(os.path.join(d, man.unis_fname(circ)))
    meas = F.load_meas_layers(os.path.join(d, man.meas_fname(p, circ)))
    return man, coev_probs(man.nqubits, unis, meas, runs, bonds, rng, exact_cap)


def coev_probs_batching(nqubits: int, unis: Sequence[Sequence[Tuple[int, np.ndarray]]],
                        meas_per_p: Sequence[Sequence[Sequence[F.Meas]]], runs: int = 32,
                        bonds: Sequence[Optional[int]] = BONDS_BATCHING,
                        rng: Optional[np.random.Generator] = None,
                        exact_cap: Optional[int] = None) -> Dict[str, np.ndarray]:
    """haar_coev_probs_batching.rs:43-70 and :139-201 for one circuit: per p
    and run, an untruncated state runs ``apply_main_layer_probs``
    (workspace/lib/haar.rs:136-161: the layer's unitaries, then each listed
    measurement in list order with its own full refactor sweep); then a
    ``Const(chi)`` state per chi replays the outcomes as post-selections.

    Returns ``traj`` (n_p x runs x depth x n, i8: +1 / -1 for outcome 1 / 0,
    0 where unmeasured) and ``prob`` (n_p x runs x n_b x depth x n, f64, -1
    where unmeasured; index 0 the untruncated state, then each chi of
    ``bonds`` in order), with ``chi`` = [0, chi...]."""
    rng = rng if rng is not None else np.random.default_rng()
    n = int(nqubits)
    depth = len(unis)
    chis = [c for c in bonds if c is not None]
    nb = len(chis) + 1
    cap = int(exact_cap) if exact_cap is not None else exact_bond_cap(n)
    eng = Engine(max([cap] + chis))
    traj = np.zeros((len(meas_per_p), runs, depth, n), dtype=np.int8)
    prob = np.full((len(meas_per_p), runs, nb, depth, n), -1.0)

    def layer_probs(st, t, layer):
        st.apply_unis(unis[t])
        outs, prs = [], []
        for m in layer:
            if isinstance(m, F.Rand):
                probs = st.probs(m.k)
                res = st.measure_prob(m.k, rng.random())
                if probs is None or res is None:
                    raise ValueError(f"measurement of qubit {m.k} out of range")
                o = res[0]
                prs.append((m.k, probs[o]))
            else:
                o = int(m.outcome)
                pr = st.prob(m.k, o)
                if pr is None:
                    raise ValueError(f"measurement of qubit {m.k} out of range")
                st.measure_postsel_prob(m.k, o)
                prs.append((m.k, pr))
            outs.append(F.Proj(m.k, o))
        return outs, prs

    for j, meas in enumerate(meas_per_p):
        if len(meas) != depth:
            raise ValueError("unitary and measurement layer counts differ")
        for r in range(runs):
            st = eng.new_qubits(n, None, chi_cap=cap)
            rec: List[List[F.Meas]] = []
            for t in range(depth):
                outs, prs = layer_probs(st, t, meas[t])
                rec.append(outs)
                for m in outs:
                    traj[j, r, t, m.k] = 1 if m.outcome else -1
                for k, pr in prs:
                    prob[j, r, 0, t, k] = pr
            del st
            for b, chi in enumerate(chis, start=1):
                st = eng.new_qubits(n, ("const", chi))
                for t in range(depth):
                    _, prs = layer_probs(st, t, rec[t])
                    for k, pr in prs:
                        prob[j, r, b, t, k] = pr
                del st
    return {"traj": traj, "prob": prob, "chi": np.asarray([0] + chis, dtype=np.int32)}


def save_coev_probs_batching(path, nqubits: int, depth: int, circ: int, circs: int, runs: int,
                             p_meas: Sequence[float], seed: int,
                             res: Dict[str, np.ndarray]) -> None:
    """The output file of haar_coev_probs_batching.rs:202-225."""
    np.savez(path,
             size=np.asarray([nqubits], dtype=np.int32),
             depth=np.asarray([depth], dtype=np.int32),
             circ=np.asarray([circ], dtype=np.int32),
             circs=np.asarray([circs], dtype=np.int32),
             runs=np.asarray([runs], dtype=np.int32),
             p_meas=np.asarray(p_meas, dtype=np.float64),
             chi=res["chi"], traj=res["traj"], prob=res["prob"],
             seed=np.asarray([seed], dtype=np.int32))
