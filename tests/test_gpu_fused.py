"""One-launch row-split Jacobi steps.  A lone SVD (or up to 4 in one launch)
whose rows are split over several workgroups per block pair runs each step
as one launch on two X buffers (every workgroup of a pair reads the whole
panel from one and writes its rotated slice to the other); a task that ran
an odd number of sweeps is copied back.  With TNMPS_SPLIT_FUSED=0 the same
steps take two launches (partial Gram matrices, then solve and apply).  Both
must compute the same SVDs to rounding, and match the oracle (zgesvd,
mps.rs:236-303) to the parity tolerances of the other GPU tests."""
import numpy as np
import pytest

from helpers import assert_spectra_equiv
from oracle.mps_oracle import run_trajectory
from tnmps import MPS, Engine
from tnmps.workload import haar_circuit

pytestmark = pytest.mark.gpu


def close(a, b, rtol=1e-10, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


# the lone refactors' QR-preconditioned Y reach 128 x 128 (rows split over 2
# workgroups per pair) and 256 x 256 (over 4)
CASES = [(18, 10, 0.15, 128, 22), (16, 10, 0.2, 256, 23)]


def _solo(n, chi, c, record, fused, monkeypatch):
    monkeypatch.setenv("TNMPS_SPLIT_FUSED", "1" if fused else "0")
    st = MPS.new_qubits(n, ("const", chi))
    st.set_timing(True)
    got = st.run_circuit(c.layers, forced=record)
    return st, got


@pytest.mark.parametrize("n,depth,p,chi,seed", CASES)
def test_fused_matches_two_launch_and_oracle(gpu, monkeypatch, n, depth, p, chi, seed):
    c = haar_circuit(n, depth, p, seed)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    a, ga = _solo(n, chi, c, ref["record"], True, monkeypatch)
    b, gb = _solo(n, chi, c, ref["record"], False, monkeypatch)
    # the one-launch steps ran: about the same sweeps in fewer launches
    assert a.perf()["jacobi_launches"] < 0.8 * b.perf()["jacobi_launches"]
    # the two forms sum the panel Gram matrices in different orders, so they
    # agree to round-off (a singular value at the rank rule's EPSILON
    # threshold may be kept by one and dropped by the other)
    assert_spectra_equiv([a.schmidt(bd) for bd in range(n - 1)],
                         [b.schmidt(bd) for bd in range(n - 1)], rtol=1e-12)
    assert_spectra_equiv([a.schmidt(bd) for bd in range(n - 1)],
                         [ref["state"].L[bd] for bd in range(n - 1)])
    for x, y, z in zip(ga["entropy"], gb["entropy"], ref["entropy"]):
        assert close(x, z) and close(x, y, rtol=1e-12)
    for k in range(n):
        for x, y in zip(a.probs(k), ref["state"].probs(k)):
            assert close(x, y)


@pytest.mark.parametrize("count", [3, 6])
def test_fused_batches_match_solo(gpu, monkeypatch, count):
    """Batches of 3 handles (row-split launches of up to 4 tasks take one
    launch per step) and of 6 (more than 4 tasks take two) against solo runs
    of each handle."""
    monkeypatch.setenv("TNMPS_SPLIT_FUSED", "1")
    n, chi, depth = 14, 64, 10
    circuits = [haar_circuit(n, depth, 0.2, 300 + i) for i in range(count)]
    refs = [run_trajectory(n, cc.layers, ("const", chi), rs=cc.draws) for cc in circuits]
    eng = Engine(chi)
    hs = [eng.new_qubits(n, ("const", chi)) for _ in range(count)]
    for t in range(depth):
        eng.batch_bilayer(hs, [cc.layers[t] for cc in circuits],
                          forced=[r["record"][t] for r in refs])
    for i in range(count):
        solo = MPS.new_qubits(n, ("const", chi))
        solo.run_circuit(circuits[i].layers, forced=refs[i]["record"])
        # (a batch of more than 4 takes the two-launch form, its solo run the
        # one-launch form: round-off apart, as above)
        assert_spectra_equiv([hs[i].schmidt(bd) for bd in range(n - 1)],
                             [solo.schmidt(bd) for bd in range(n - 1)], rtol=1e-12)
        assert_spectra_equiv([hs[i].schmidt(bd) for bd in range(n - 1)],
                             [refs[i]["state"].L[bd] for bd in range(n - 1)])
