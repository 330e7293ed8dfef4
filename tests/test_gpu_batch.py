"""Lock-step batching of many trajectories on one engine, the FP32
contraction mode, and the Jacobi launch shapes.

Each trajectory of a batch must give the oracle's results (same tolerances
as tests/test_gpu_parity.py) and the same results as when run alone.
"""
import os

import numpy as np
import pytest

from helpers import assert_spectrum_close
from oracle.mps_oracle import run_trajectory
from tnmps import MPS, Engine
from tnmps.workload import haar_circuit

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-10, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


def _specs():
    # (n, depth, p, chi, seed): uneven sizes and bond dimensions in one batch
    return [(8, 8, 0.2, 16, 1), (12, 10, 0.1, 32, 2), (10, 10, 0.3, 8, 3),
            (16, 8, 0.05, 32, 4), (6, 6, 0.5, 4, 5), (12, 12, 0.15, 16, 6)]


def test_batch_matches_oracle_and_solo(gpu):
    specs = _specs()
    circuits = [haar_circuit(n, d, p, s) for (n, d, p, _, s) in specs]
    refs = [run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
            for (n, _, _, chi, _), c in zip(specs, circuits)]
    eng = Engine(32)
    states = [eng.new_qubits(n, ("const", chi)) for (n, _, _, chi, _) in specs]
    depth = max(d for (_, d, _, _, _) in specs)
    recs = [[] for _ in specs]
    probs = [[] for _ in specs]
    ents = [[] for _ in specs]
    for t in range(depth):
        idx = [i for i, (_, d, _, _, _) in enumerate(specs) if t < d]
        res = eng.batch_bilayer([states[i] for i in idx],
                                [circuits[i].layers[t] for i in idx],
                                forced=[refs[i]["record"][t] for i in idx])
        for i, (rec, prs, ent) in zip(idx, res):
            recs[i].append(rec)
            probs[i].append(prs)
            ents[i].append(ent)
    for i, (n, d, p, chi, s) in enumerate(specs):
        ref = refs[i]
        assert recs[i] == ref["record"]
        for a, b in zip(ents[i], ref["entropy"][1:]):
            assert close(a, b), f"traj {i} entropy {a} vs {b}"
        for la, lb in zip(probs[i], ref["probs"]):
            for (ka, pa), (kb, pb) in zip(la, lb):
                assert ka == kb and close(pa, pb)
        assert states[i].bond_dims() == ref["state"].bond_dims()
        for b in range(n - 1):
            assert_spectrum_close(states[i].schmidt(b), ref["state"].L[b], msg=f"traj {i} bond {b}")
        # the same trajectory alone, on its own engine
        solo = MPS.new_qubits(n, ("const", chi))
        got = solo.run_circuit(circuits[i].layers, forced=ref["record"])
        for b in range(n - 1):
            assert_spectrum_close(solo.schmidt(b), states[i].schmidt(b), rtol=1e-12,
                                  msg=f"solo vs batch traj {i} bond {b}")
        for a, c in zip(got["entropy"][1:], ents[i]):
            assert close(a, c, rtol=1e-12)


def test_batch_sampled_draws(gpu):
    """Meas::Rand in a batch consumes each trajectory's own draws."""
    specs = [(10, 6, 0.3, 16, 11), (8, 6, 0.3, 8, 12)]
    circuits = [haar_circuit(n, d, p, s) for (n, d, p, _, s) in specs]
    refs = [run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
            for (n, _, _, chi, _), c in zip(specs, circuits)]
    eng = Engine(16)
    states = [eng.new_qubits(n, ("const", chi)) for (n, _, _, chi, _) in specs]
    its = [iter(c.draws) for c in circuits]
    recs = [[], []]
    for t in range(6):
        res = eng.batch_bilayer(states, [c.layers[t] for c in circuits], draws=its)
        for i in range(2):
            recs[i].append(res[i][0])
    assert recs[0] == refs[0]["record"] and recs[1] == refs[1]["record"]


def test_fp32_contraction_error(gpu):
    """TN_PREC_F32: theta in FP32 with FP32 accumulation, SVD in FP64.  The
    entropy error against the FP64 oracle stays within 1e-6 after every one
    of the 12 layers, and so do the final Schmidt values."""
    n, depth, p, chi = 12, 12, 0.1, 16
    c = haar_circuit(n, depth, p, 77)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    st = MPS.new_qubits(n, ("const", chi)).set_precision("f32")
    got = st.run_circuit(c.layers, forced=ref["record"])
    errs = [abs(a - b) for a, b in zip(got["entropy"], ref["entropy"])]
    print(f"FP32 n={n} chi={chi}: entropy errors per layer", [f"{e:.1e}" for e in errs])
    assert max(errs) <= 1e-6, errs  # every layer (measured: at most 4.3e-7)
    assert max(errs) > 0.0  # the FP32 path really ran
    s0 = max(float(np.max(np.abs(st.schmidt(b)[: len(ref["state"].L[b])]
                                 - ref["state"].L[b][: len(st.schmidt(b))])))
             for b in range(n - 1))
    assert s0 <= 1e-6


def test_step_and_row_split_jacobi_agree(gpu, monkeypatch):
    """The one-workgroup-per-pair step (TNMPS_SPLIT_MAX=1) and the row-split
    step (the default for lone updates) give the oracle's spectra, and so do
    the wide 16-column panels (TNMPS_BW=16), and the row split forced down to
    64-row slices (TNMPS_SPLIT_ROWS=64)."""
    n, depth, p, chi = 12, 10, 0.1, 32
    c = haar_circuit(n, depth, p, 31)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    runs = []
    for env in ({}, {"TNMPS_SPLIT_MAX": "1"}, {"TNMPS_BW": "16"},
                {"TNMPS_BW": "16", "TNMPS_SPLIT_MAX": "1"}, {"TNMPS_SPLIT_ROWS": "64"}):
        for k in ("TNMPS_SPLIT_MAX", "TNMPS_BW", "TNMPS_SPLIT_ROWS"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        st = MPS.new_qubits(n, ("const", chi))
        st.run_circuit(c.layers, forced=ref["record"])
        runs.append(st)
    for st in runs:
        assert st.bond_dims() == ref["state"].bond_dims()
        for bd in range(n - 1):
            assert_spectrum_close(st.schmidt(bd), ref["state"].L[bd])


def test_row_split_matches_oracle(gpu):
    """Bonds at chi = 64 (theta 128 x 128): lone refactors take the row-split
    path at its default slice size; spectra, records and entropies match the
    oracle."""
    n, depth, p, chi = 16, 12, 0.15, 64
    c = haar_circuit(n, depth, p, 77)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    st = MPS.new_qubits(n, ("const", chi))
    got = st.run_circuit(c.layers, forced=ref["record"])
    assert st.bond_dims() == ref["state"].bond_dims()
    for bd in range(n - 1):
        assert_spectrum_close(st.schmidt(bd), ref["state"].L[bd])
    for a, b in zip(got["entropy"], ref["entropy"]):
        assert abs(a - b) <= 1e-12 + 1e-10 * abs(b)


def test_active_task_list_bitwise(gpu, monkeypatch):
    """Large lock-step batches run their Jacobi steps over a per-sweep list of
    unconverged tasks (batches of TNMPS_LIST_MIN tasks and up).  Off, default,
    and forced on every batch, the states are bitwise the same, and match the
    oracle."""
    T, n, depth, chi = 40, 10, 8, 16
    ps = [0.05 + 0.3 * (i % 7) / 6 for i in range(T)]
    circuits = [haar_circuit(n, depth, ps[i], 500 + i) for i in range(T)]
    refs = [run_trajectory(n, c.layers, ("const", chi), rs=c.draws) for c in circuits[:4]]
    recs = [r["record"] for r in refs]
    for c in circuits[4:]:
        st = MPS.new_qubits(n, ("const", chi))
        recs.append(st.run_circuit(c.layers, rs=c.draws)["record"])
    runs = []
    for env in ({"TNMPS_LIST_MIN": "100000"}, {},
                {"TNMPS_LIST_MIN": "1"}):
        monkeypatch.delenv("TNMPS_LIST_MIN", raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        eng = Engine(chi)
        states = [eng.new_qubits(n, ("const", chi)) for _ in range(T)]
        for t in range(depth):
            eng.batch_bilayer(states, [c.layers[t] for c in circuits],
                              forced=[r[t] for r in recs])
        runs.append([([st.gamma(k) for k in range(n)], [st.schmidt(b) for b in range(n - 1)])
                     for st in states])
    for other in runs[1:]:
        for (g0, l0), (g1, l1) in zip(runs[0], other):
            assert all(np.array_equal(a, b) for a, b in zip(g0, g1))
            assert all(np.array_equal(a, b) for a, b in zip(l0, l1))
    for i, ref in enumerate(refs):
        for b in range(n - 1):
            assert_spectrum_close(runs[1][i][1][b], ref["state"].L[b], msg=f"traj {i} bond {b}")


def test_engines_in_threads_bitwise(gpu):
    """Lock-step groups on separate engines (streams), driven from separate
    host threads at once, give bitwise the states that the same groups give
    one after the other."""
    import threading

    T, n, depth, chi = 12, 10, 6, 16
    circuits = [haar_circuit(n, depth, 0.1 + 0.05 * (i % 4), 900 + i) for i in range(T)]
    groups = [list(range(g, T, 3)) for g in range(3)]

    def run(g, out):
        eng = Engine(chi)
        sts = [eng.new_qubits(n, ("const", chi)) for _ in groups[g]]
        its = [iter(circuits[i].draws) for i in groups[g]]
        for t in range(depth):
            eng.batch_bilayer(sts, [circuits[i].layers[t] for i in groups[g]], draws=its)
        eng.sync()
        out[g] = [[st.schmidt(b) for b in range(n - 1)] + [st.gamma(k) for k in range(n)]
                  for st in sts]

    seq = {}
    for g in range(3):
        run(g, seq)
    par = {}
    th = [threading.Thread(target=run, args=(g, par)) for g in range(3)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for g in range(3):
        for a, b in zip(seq[g], par[g]):
            assert all(np.array_equal(u, v) for u, v in zip(a, b))
