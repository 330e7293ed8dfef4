"""The 16 x 16 panel solves' schedule of pairs inside each 8-column block.

Every step of a sweep rotates the pairs across the two blocks of each panel;
the pairs inside a block are rotated on every TNMPS_INNER_EVERY-th step only
(8 by default, 1 = every step).  Both schedules converge to the same SVDs:
a trajectory run both ways must give the same spectra, entropies and
probabilities to round-off, both matching the oracle (zgesvd,
mps.rs:236-303), in about as many sweeps."""
import numpy as np
import pytest

from helpers import assert_spectra_equiv
from oracle.mps_oracle import run_trajectory
from tnmps import MPS
from tnmps.workload import haar_circuit

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-10, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


# lone refactors with Y of 128 and 256 rows (16 and 32 blocks: the sparse
# schedule rotates inside blocks on 2 of 15 and 4 of 31 steps)
CASES = [(18, 10, 0.15, 128, 22), (16, 10, 0.2, 256, 23)]


def _run(n, chi, c, record, every, monkeypatch):
    monkeypatch.setenv("TNMPS_INNER_EVERY", str(every))
    st = MPS.new_qubits(n, ("const", chi))
    st.set_timing(True)
    got = st.run_circuit(c.layers, forced=record)
    return st, got


@pytest.mark.parametrize("n,depth,p,chi,seed", CASES)
def test_inner_schedule_matches_every_step_and_oracle(gpu, monkeypatch, n, depth, p, chi, seed):
    c = haar_circuit(n, depth, p, seed)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    a, ga = _run(n, chi, c, ref["record"], 8, monkeypatch)
    b, gb = _run(n, chi, c, ref["record"], 1, monkeypatch)
    # about as many sweeps (the CPU emulation of both schedules: 146 against
    # 144 sweeps on 37 captured inputs)
    assert a.perf()["sweeps"] <= 1.1 * b.perf()["sweeps"] + 2
    sa = [a.schmidt(bd) for bd in range(n - 1)]
    sb = [b.schmidt(bd) for bd in range(n - 1)]
    assert_spectra_equiv(sa, sb, rtol=1e-11)
    assert_spectra_equiv(sa, [ref["state"].L[bd] for bd in range(n - 1)])
    for x, y, z in zip(ga["entropy"], gb["entropy"], ref["entropy"]):
        assert close(x, z) and close(x, y, rtol=1e-11)
    for k in range(n):
        for x, y in zip(a.probs(k), ref["state"].probs(k)):
            assert close(x, y)
