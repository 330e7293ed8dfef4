"""SVD task pool.  With TNMPS_POOL_MIN set, from that many handles on a
batch's measurement schedules run as a pool of SVD tasks: each handle keeps at most
one refactor in flight, every iteration runs one Jacobi sweep over all
tasks in flight, and a converged task is finished and replaced at once by
its handle's next refactor.  Every handle must give what it gives alone and
in a lock-step batch (to rounding, 1e-12) and what the oracle gives
(apply_meas_prob_layer, mps.rs:2206-2252, to the parity tolerances)."""
import numpy as np
import pytest

from helpers import assert_spectra_equiv, assert_spectrum_close
from oracle.mps_oracle import run_trajectory
from tnmps import MPS, Engine
from tnmps.workload import haar_circuit

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-10, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


# (n, depth, p, chi, seed): uneven sizes, bond caps and measurement rates
SPECS = [(8, 8, 0.2, 16, 1), (12, 10, 0.1, 32, 2), (10, 10, 0.3, 8, 3), (16, 8, 0.05, 32, 4),
         (6, 6, 0.5, 4, 5), (12, 12, 0.15, 16, 6), (14, 9, 0.25, 24, 7), (9, 11, 0.4, 12, 8)]


def _run_batch(specs, circuits, records, chi_eng):
    eng = Engine(chi_eng)
    states = [eng.new_qubits(n, ("const", chi)) for (n, _, _, chi, _) in specs]
    depth = max(d for (_, d, _, _, _) in specs)
    recs = [[] for _ in specs]
    probs = [[] for _ in specs]
    ents = [[] for _ in specs]
    for t in range(depth):
        idx = [i for i, (_, d, _, _, _) in enumerate(specs) if t < d]
        res = eng.batch_bilayer([states[i] for i in idx], [circuits[i].layers[t] for i in idx],
                                forced=[records[i][t] for i in idx])
        for i, (rec, prs, ent) in zip(idx, res):
            recs[i].append(rec)
            probs[i].append(prs)
            ents[i].append(ent)
    return states, recs, probs, ents


def test_pool_matches_lockstep_solo_and_oracle(gpu, monkeypatch):
    circuits = [haar_circuit(n, d, p, s) for (n, d, p, _, s) in SPECS]
    refs = [run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
            for (n, _, _, chi, _), c in zip(SPECS, circuits)]
    records = [r["record"] for r in refs]
    monkeypatch.setenv("TNMPS_POOL_MIN", "2")
    pool = _run_batch(SPECS, circuits, records, 32)
    monkeypatch.setenv("TNMPS_POOL_MIN", "1000")
    lock = _run_batch(SPECS, circuits, records, 32)
    for i, (n, d, p, chi, s) in enumerate(SPECS):
        st, ref = pool[0][i], refs[i]
        assert pool[1][i] == ref["record"] == lock[1][i]
        for a, b, c in zip(pool[3][i], ref["entropy"][1:], lock[3][i]):
            assert close(a, b), f"traj {i} entropy {a} vs oracle {b}"
            assert close(a, c, rtol=1e-12), f"traj {i} entropy {a} vs lock-step {c}"
        for la, lb in zip(pool[2][i], ref["probs"]):
            for (ka, pa), (kb, pb) in zip(la, lb):
                assert ka == kb and close(pa, pb)
        assert_spectra_equiv([st.schmidt(b) for b in range(n - 1)],
                             [ref["state"].L[b] for b in range(n - 1)])
        for b in range(n - 1):
            assert_spectrum_close(st.schmidt(b), lock[0][i].schmidt(b), rtol=1e-12,
                                  msg=f"pool vs lock-step traj {i} bond {b}")
        solo = MPS.new_qubits(n, ("const", chi))
        got = solo.run_circuit(circuits[i].layers, forced=records[i])
        for a, c in zip(got["entropy"][1:], pool[3][i]):
            assert close(a, c, rtol=1e-12)


def test_pool_mixed_precisions(gpu, monkeypatch):
    """20 handles in a pool (TNMPS_POOL_MIN=16): FP64 handles equal their
    solo runs to 1e-12 and FP32 ones theirs to the same, with tasks of both
    precisions in flight together."""
    monkeypatch.setenv("TNMPS_POOL_MIN", "16")
    n, depth, chi = 12, 8, 16
    circuits = [haar_circuit(n, depth, 0.2, 100 + i) for i in range(20)]
    refs = [run_trajectory(n, c.layers, ("const", chi), rs=c.draws) for c in circuits]
    eng = Engine(chi)
    states = [eng.new_qubits(n, ("const", chi)) for _ in range(20)]
    for i in range(0, 20, 3):
        states[i].set_precision("f32")
    ents = [[] for _ in range(20)]
    for t in range(depth):
        res = eng.batch_bilayer(states, [c.layers[t] for c in circuits],
                                forced=[r["record"][t] for r in refs])
        for i, (_, _, e) in enumerate(res):
            ents[i].append(e)
    for i in range(20):
        solo = MPS.new_qubits(n, ("const", chi))
        if i % 3 == 0:
            solo.set_precision("f32")
        got = solo.run_circuit(circuits[i].layers, forced=refs[i]["record"])
        for a, c in zip(got["entropy"][1:], ents[i]):
            assert close(a, c, rtol=1e-12), (i, a, c)
        if i % 3:
            for a, b in zip(ents[i], refs[i]["entropy"][1:]):
                assert close(a, b)
