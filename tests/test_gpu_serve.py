"""Continuous batching: every request served in the mix (staggered arrivals,
different n, p and chi_max, joining and leaving the lock-step batch) gives
the same outcome record, probabilities and entropy curve as the same
trajectory run alone, and the oracle's (Meas::Rand with the same draws)."""
import numpy as np
import pytest

from oracle.mps_oracle import run_trajectory
from tnmps import MPS, Engine
from tnmps.serve import Request, Server, mps_bytes, workspace_bytes
from tnmps.workload import haar_circuit

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-10, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


def _requests():
    specs = [(8, 0.2, 16, 0.0), (12, 0.1, 32, 0.0), (10, 0.3, 8, 1.0), (6, 0.0, 4, 2.5),
             (14, 0.15, 32, 3.0), (8, 0.4, 16, 7.0)]
    reqs = []
    for i, (n, p, chi, t) in enumerate(specs):
        c = haar_circuit(n, n, p, 500 + i)
        reqs.append(Request(rid=i, n=n, p=p, chi=chi, layers=c.layers, draws=c.draws, arrival=t))
    return reqs


def test_serving_mix_equals_alone(gpu):
    reqs = _requests()
    # a budget that holds the engine's workspace (capped at 8 slots, so
    # batches are chunked too) and only part of the stream forces queueing
    # and admission
    ws_cap = 8 * (16.0 * 2 * 64 * 64 + 64 * 64 + 4096)
    budget = workspace_bytes(32, ws_cap) + sum(mps_bytes(r.n, r.chi) for r in reqs[:3])
    srv = Server(chi_max=32, hbm_budget=budget, ws_cap=ws_cap)
    assert srv.ws_held == workspace_bytes(32, ws_cap)
    # simulated clock: one time unit per tick, so arrivals interleave with layers
    ticks = {"t": 0.0}

    def clock():
        ticks["t"] += 0.5
        return ticks["t"]

    srv.run(reqs, clock=clock, realtime=False)
    assert srv.ticks > 0
    for r in reqs:
        assert r.done is not None and r.admitted is not None and r.admitted >= r.arrival
        # alone on its own engine
        solo = MPS.new_qubits(r.n, ("const", r.chi))
        got = solo.run_circuit(r.layers, rs=iter(r.draws))
        assert got["record"] == r.record
        for a, b in zip(got["entropy"], r.entropy):
            assert close(a, b, rtol=1e-12), (r.rid, a, b)
        for la, lb in zip(got["probs"], r.probs):
            for (ka, pa), (kb, pb) in zip(la, lb):
                assert ka == kb and close(pa, pb, rtol=1e-12)
        # and the oracle
        ref = run_trajectory(r.n, r.layers, ("const", r.chi), rs=r.draws)
        assert ref["record"] == r.record
        for a, b in zip(ref["entropy"], r.entropy):
            assert close(a, b), (r.rid, a, b)
    # admission respected the HBM budget
    assert srv.peak_held <= budget


def test_serving_generated_equals_alone(gpu):
    """Requests generated on the device (seed, trajectory id) served in a mix
    with carried circuits give the same results as the same generated
    trajectory run alone."""
    reqs = _requests()[:3]
    seed = 4242
    for i, (n, p, chi, t) in enumerate([(10, 0.2, 16, 0.0), (12, 0.3, 8, 0.5), (8, 0.1, 32, 2.0)]):
        reqs.append(Request(rid=10 + i, n=n, p=p, chi=chi, arrival=t, gen_seed=seed,
                            traj_id=100 + i, gen_depth=n))
    srv = Server(chi_max=32)
    ticks = {"t": 0.0}

    def clock():
        ticks["t"] += 0.5
        return ticks["t"]

    srv.run(reqs, clock=clock, realtime=False)
    for r in reqs[3:]:
        assert r.done is not None and len(r.record) == r.depth
        eng = Engine(r.chi)
        st = eng.new_qubits(r.n, ("const", r.chi))
        ent = [0.0]
        for t in range(r.depth):
            (rec, prs, e), = eng.batch_bilayer_gen([st], seed, [r.traj_id], [t], [r.p])
            assert rec == r.record[t]
            for (ka, pa), (kb, pb) in zip(prs, r.probs[t]):
                assert ka == kb and close(pa, pb, rtol=1e-12)
            ent.append(e)
        for a, b in zip(ent, r.entropy):
            assert close(a, b, rtol=1e-12)


@pytest.mark.parametrize("dispatch", ["least_loaded", "round_robin"])
def test_serving_workers_equal_alone(gpu, dispatch):
    """Two lock-step workers (engines and host threads) sharing one budget:
    every request, carried or generated on the device, gives the results it
    gives alone, whichever rule deals the requests to the workers."""
    reqs = _requests()
    for r in reqs:
        r.arrival *= 0.05   # real time: arrivals within 0.4 s
    seed = 777
    for i, (n, p, chi, t) in enumerate([(10, 0.2, 16, 0.0), (12, 0.3, 8, 0.1)]):
        reqs.append(Request(rid=20 + i, n=n, p=p, chi=chi, arrival=t, gen_seed=seed,
                            traj_id=200 + i, gen_depth=n))
    ws_cap = 16 * (16.0 * 2 * 64 * 64 + 64 * 64 + 4096)
    budget = 2 * workspace_bytes(32, ws_cap) + sum(mps_bytes(r.n, r.chi) for r in reqs[:4])
    srv = Server(chi_max=32, hbm_budget=budget, workers=2, dispatch=dispatch, ws_cap=ws_cap)
    srv.run(reqs)
    assert srv.peak_held <= budget
    assert srv.stats()[0] > 0
    for r in reqs:
        assert r.done is not None and len(r.record) == r.depth
        if r.gen_seed is None:
            solo = MPS.new_qubits(r.n, ("const", r.chi))
            got = solo.run_circuit(r.layers, rs=iter(r.draws))
            assert got["record"] == r.record
            for a, b in zip(got["entropy"], r.entropy):
                assert close(a, b, rtol=1e-12), (r.rid, a, b)
        else:
            eng = Engine(r.chi)
            st = eng.new_qubits(r.n, ("const", r.chi))
            for t in range(r.depth):
                (rec, prs, e), = eng.batch_bilayer_gen([st], seed, [r.traj_id], [t], [r.p])
                assert rec == r.record[t]
                assert close(e, r.entropy[t + 1], rtol=1e-12)
