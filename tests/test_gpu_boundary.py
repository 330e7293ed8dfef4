"""The C ABI's failure reporting and the multi-operator entry point, on the GPU.

- TN_CAPACITY: a None / Cutoff state whose rank rule asks for more values
  than its bond allocation keeps that many and the call says so (the
  reference, mps.rs:157-176, has no bound).
- TN_NO_CONVERGENCE: an SVD stopped at the sweep limit.  The condition is
  forced from the host with a sweep limit of 1 (TNMPS_MAX_SWEEPS, read when
  an engine is made); every kernel still runs to its normal end.
- apply_unitary2_multi (mps.rs:1393-1411) against the oracle.
"""
import numpy as np
import pytest

from helpers import assert_spectrum_close
from oracle.mps_oracle import VidalMPS
from tnmps import MPS, MPSError
from tnmps import _lib as L
from tnmps.workload import haar4

pytestmark = pytest.mark.gpu


def _brickwork_until(st, n, rng, layers, catch):
    """Haar brickwork gates; returns the first MPSError of kind `catch`."""
    for t in range(layers):
        for k in range(t % 2, n - 1, 2):
            try:
                st.apply_unitary2(k, haar4(rng))
            except MPSError as exc:
                if exc.kind == catch:
                    return exc
                raise
    return None


def test_capacity_reported_for_cutoff(gpu):
    """Cutoff(1e-14) at N=12 with chi_cap=8: the exact bonds want up to 64
    values; the first update that wants more than 8 returns TN_CAPACITY,
    keeps 8 normalised values and leaves a valid state."""
    n = 12
    st = MPS.new_qubits(n, ("cutoff", 1e-14), chi_cap=8)
    exc = _brickwork_until(st, n, np.random.default_rng(5), 8, "Capacity")
    assert exc is not None and exc.code == L.TN_CAPACITY
    assert st.take_flags() & L.TN_FLAG_CAPACITY
    assert st.take_flags() == 0  # taken
    assert st.max_bond_dim() == 8
    for b in range(n - 1):
        lam = st.schmidt(b)
        assert abs(np.sum(lam ** 2) - 1.0) < 1e-12
    for k in range(n):
        assert abs(sum(st.probs(k)) - 1.0) < 1e-12


def test_no_capacity_within_bound(gpu):
    """None at N=10 has exact bonds up to 32: the default bound holds them,
    so no call reports a clamp and the flags stay clear."""
    n = 10
    st = MPS.new_qubits(n, None)
    assert _brickwork_until(st, n, np.random.default_rng(6), 12, "Capacity") is None
    assert st.take_flags() == 0
    assert st.max_bond_dim() == 32


def test_no_convergence_reported(gpu, monkeypatch):
    """A sweep limit of 1 stops an entangling gate's SVD before its
    tolerance: the call returns TN_NO_CONVERGENCE, the handle keeps the flag,
    and the state is still written (finite, unit-norm spectra)."""
    monkeypatch.setenv("TNMPS_MAX_SWEEPS", "1")
    n = 8
    st = MPS.new_qubits(n, ("const", 16))
    exc = _brickwork_until(st, n, np.random.default_rng(7), 6, "NoConvergence")
    assert exc is not None and exc.code == L.TN_NO_CONVERGENCE
    assert st.take_flags() & L.TN_FLAG_NO_CONVERGENCE
    for b in range(n - 1):
        lam = st.schmidt(b)
        assert np.all(np.isfinite(lam))
        assert abs(np.sum(lam ** 2) - 1.0) < 1e-12


def test_default_sweep_limit_converges(gpu):
    """The same gates with the default limit report nothing."""
    n = 8
    st = MPS.new_qubits(n, ("const", 16))
    assert _brickwork_until(st, n, np.random.default_rng(7), 6, "NoConvergence") is None
    assert st.take_flags() == 0


@pytest.mark.parametrize("count", [0, 1, 3])
def test_apply_unitary2_multi_vs_oracle(gpu, count):
    """The product op[count-1] ... op[0] applied as one apply_unitary2
    (mps.rs:1404-1411), on an entangled state; out-of-range k is a no-op."""
    n = 8
    rng = np.random.default_rng(40 + count)
    st = MPS.new_qubits(n, None)
    ora = VidalMPS(n, None)
    for t in range(4):
        for k in range(t % 2, n - 1, 2):
            u = haar4(rng)
            st.apply_unitary2(k, u)
            ora.apply_unitary2(k, u)
    ops = [haar4(rng) for _ in range(count)]
    prod = np.eye(4, dtype=np.complex128)
    for op in ops:
        prod = op @ prod
    st.apply_unitary2_multi(3, ops)
    ora.apply_unitary2(3, prod)
    assert st.bond_dims() == ora.bond_dims()
    for b in range(n - 1):
        assert_spectrum_close(st.schmidt(b), ora.L[b], rtol=1e-10)
    for k in range(n):
        np.testing.assert_allclose(st.probs(k), ora.probs(k), atol=1e-12)
    before = [st.schmidt(b) for b in range(n - 1)]
    st.apply_unitary2_multi(n - 1, ops)  # k >= n - 1: no-op (mps.rs:1400)
    for b in range(n - 1):
        np.testing.assert_array_equal(st.schmidt(b), before[b])
