"""Factored updates.  A refactor of bond k whose site k was last written as
the left factor of an update (lambda[k-1] Gamma[k] has orthonormal columns),
or whose site k+1 was last written as the right factor (Gamma[k+1]
lambda[k+1] has orthonormal rows), decomposes only the other factor of
theta.  Results must match the oracle (zgesvd of the full theta,
mps.rs:846-880) and the full-theta path of this library (TNMPS_FACTORED=0)
to the parity tolerances of the other GPU tests."""
import numpy as np
import pytest

from helpers import assert_spectrum_close
from oracle.mps_oracle import run_trajectory
from tnmps import MPS
from tnmps.workload import haar_circuit

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def close(a, b, rtol=RTOL, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


# bonds saturate at chi within a few layers, so most refactors are factored
CASES = [(24, 16, 0.1, 16, 6), (32, 12, 0.1, 8, 7), (16, 14, 0.25, 8, 9)]


def _run(n, c, chi, record):
    st = MPS.new_qubits(n, ("const", chi))
    st.set_timing(True)
    got = st.run_circuit(c.layers, forced=record)
    return st, got


@pytest.mark.parametrize("n,depth,p,chi,seed", CASES)
def test_factored_matches_oracle_and_full(gpu, monkeypatch, n, depth, p, chi, seed):
    c = haar_circuit(n, depth, p, seed)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    monkeypatch.setenv("TNMPS_FACTORED", "1")
    a, ga = _run(n, c, chi, ref["record"])
    pa = a.perf()
    assert pa["factored"] > 0, "no factored update ran"
    monkeypatch.setenv("TNMPS_FACTORED", "0")
    b, gb = _run(n, c, chi, ref["record"])
    assert b.perf()["factored"] == 0
    assert a.bond_dims() == ref["state"].bond_dims() == b.bond_dims()
    for bd in range(n - 1):
        assert_spectrum_close(a.schmidt(bd), ref["state"].L[bd], msg=f"factored bond {bd}")
        assert_spectrum_close(a.schmidt(bd), b.schmidt(bd), msg=f"factored vs full bond {bd}")
    for x, y, z in zip(ga["entropy"], gb["entropy"], ref["entropy"]):
        assert close(x, z) and close(x, y)
    for la, lr in zip(ga["probs"], ref["probs"]):
        for (ka, pa_), (kr, pr) in zip(la, lr):
            assert ka == kr and close(pa_, pr)
    for k in range(n):
        for x, y in zip(a.probs(k), ref["state"].probs(k)):
            assert close(x, y)


def test_factored_after_one_site_ops_and_sweeps(gpu):
    """One-site operators (not necessarily unitary) and projections clear the
    freshness of their site; measure_prob's refactor sweep and gates restore
    it.  Mixed sequence against the oracle."""
    from oracle.mps_oracle import VidalMPS
    from tnmps.workload import haar4
    rng = np.random.default_rng(42)
    n, chi = 14, 8
    a = MPS.new_qubits(n, ("const", chi))
    o = VidalMPS(n, ("const", chi))
    for layer in range(8):
        for k in range(layer % 2, n - 1, 2):
            u = haar4(rng)
            a.apply_unitary2(k, u)
            o.apply_unitary2(k, u)
        k1 = int(rng.integers(n))
        op = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))  # not unitary
        a.apply_unitary1(k1, op)
        o.apply_unitary1(k1, op)
        km = int(rng.integers(n))
        pa = a.measure_postsel_prob(km, layer % 2)
        po = o.measure_postsel_prob(km, layer % 2)
        assert close(pa, po), (layer, pa, po)
    assert a.bond_dims() == o.bond_dims()
    for bd in range(n - 1):
        assert_spectrum_close(a.schmidt(bd), o.L[bd], msg=f"bond {bd}")
