"""QR-preconditioned SVD.  With the tuning field ``qr`` set (the default), an
update batch whose SVDs have at most 512 rows and at least 32 columns sorts
theta's columns by norm, factors it with blocked Householder QR and runs the
Jacobi sweeps on R^H.  The result is the same SVD to rounding: Schmidt
values, probabilities and entropies must match the oracle (zgesvd of theta,
mps.rs:236-303) and the plain Jacobi path (qr = 0) to the parity tolerances
of the other GPU tests.  The cases put partial Householder panels (C not a
multiple of 16) and factored updates (C = K < R) under test."""
import numpy as np
import pytest

from helpers import assert_spectra_equiv
from oracle.mps_oracle import VidalMPS, run_trajectory
from tnmps import MPS, Engine
from tnmps.workload import haar4, haar_circuit

pytestmark = pytest.mark.gpu

RTOL = 1e-10
BUCKETS = (16, 32, 64, 128, 256, 512, 1024)


def close(a, b, rtol=RTOL, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


def table(qr):
    return [dict(c_max=c, theta_tile=32 if c <= 128 else 64, bw=8, split_max=8, split_rows=128,
                 ahead=2, qr=qr) for c in BUCKETS]


# (n, depth, p, chi, seed): C up to 40, 48, 80 and 128 columns
CASES = [(14, 10, 0.2, 20, 3), (16, 10, 0.15, 24, 5), (18, 9, 0.1, 40, 11),
         (20, 8, 0.1, 64, 2)]


@pytest.mark.parametrize("n,depth,p,chi,seed", CASES)
def test_qr_matches_oracle_and_plain(gpu, n, depth, p, chi, seed):
    c = haar_circuit(n, depth, p, seed)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    runs = {}
    for qr in (1, 0):
        st = MPS.new_qubits(n, ("const", chi))
        st.set_tuning(table(qr))
        got = st.run_circuit(c.layers, forced=ref["record"])
        runs[qr] = (st, got, st.perf()["qr_tasks"])
    (a, ga, qa), (b, gb, qb) = runs[1], runs[0]
    assert qa > 0, "the QR path did not run"
    assert qb == 0
    sa = [a.schmidt(bd) for bd in range(n - 1)]
    sb = [b.schmidt(bd) for bd in range(n - 1)]
    so = [ref["state"].L[bd] for bd in range(n - 1)]
    assert_spectra_equiv(sa, so)
    assert_spectra_equiv(sa, sb)
    for x, y, z in zip(ga["entropy"], gb["entropy"], ref["entropy"]):
        assert close(x, z) and close(x, y)
    for la, lr in zip(ga["probs"], ref["probs"]):
        for (ka, pa_), (kr, pr) in zip(la, lr):
            assert ka == kr and close(pa_, pr)
    for k in range(n):
        for x, y in zip(a.probs(k), ref["state"].probs(k)):
            assert close(x, y)


def test_qr_uneven_batch_matches_solo(gpu):
    """Handles of different chain lengths and bond caps in one engine: their
    updates share QR launches (tasks with fewer columns skip the later
    panels) and each must equal its solo run and the oracle."""
    specs = [(12, 16), (14, 24), (16, 40)]
    rng = np.random.default_rng(7)
    eng = Engine(40)
    eng.set_tuning(table(1))
    hs = [eng.new_qubits(n, ("const", chi)) for n, chi in specs]
    solo = [MPS.new_qubits(n, ("const", chi)) for n, chi in specs]
    ors = [VidalMPS(n, ("const", chi)) for n, chi in specs]
    for s in solo:
        s.set_tuning(table(1))
    for layer in range(8):
        gates = [[(k, haar4(rng)) for k in range(layer % 2, n - 1, 2)] for n, _ in specs]
        eng.batch_bilayer(hs, [(gl, []) for gl in gates])
        for s, o, gl in zip(solo, ors, gates):
            s.run_circuit([(gl, [])])
            for k, u in gl:
                o.apply_unitary2(k, u)
    assert eng.perf()["qr_tasks"] > 0
    for (n, _), h, s, o in zip(specs, hs, solo, ors):
        for bd in range(n - 1):
            np.testing.assert_allclose(h.schmidt(bd), s.schmidt(bd), rtol=0, atol=1e-12)
        assert_spectra_equiv([h.schmidt(bd) for bd in range(n - 1)],
                             [o.L[bd] for bd in range(n - 1)])
