"""MPS::from_vector (mps.rs:516-546 -> factorize, mps.rs:448-508) on the
device, against the oracle's restatement: bond dimensions, Schmidt values
(|d| <= 1e-12 + 1e-10 s0), probabilities and entropies, and the contracted
state against the normalised input vector.  The singular vectors' phases
are a gauge, so Gamma itself is not compared."""
import warnings

import numpy as np
import pytest

from helpers import assert_spectrum_close, state_overlap_err
from oracle.mps_oracle import from_vector
from tnmps import MPS, MPSError, MPSWarning
from tnmps.workload import haar4

pytestmark = pytest.mark.gpu


def _vector(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)) * 2.5


def _compare(st: MPS, o, n, psi):
    assert st.bond_dims() == o.bond_dims()
    for b in range(n - 1):
        assert_spectrum_close(st.schmidt(b), o.L[b], msg=f"bond {b}")
        e_g, e_o = st.entropy_vn(b), o.entropy_vn(b)
        assert abs(e_g - e_o) <= 1e-12 + 1e-10 * abs(e_o)
    for k in range(n):
        for a, c in zip(st.probs(k), o.probs(k)):
            assert abs(a - c) <= 1e-12 + 1e-10 * abs(c)


@pytest.mark.parametrize("n,trunc,driver", [(1, None, "gesvd"), (2, None, "gesvd"),
                                            (10, None, "gesvd"), (14, ("const", 16), "gesvd"),
                                            (12, ("cutoff", 1e-3), "gesvd"),
                                            (18, None, "gesdd")])
def test_from_vector_vs_oracle(gpu, n, trunc, driver):
    """Untruncated up to n=18 (a 512 x 512 SVD at the middle bond, 2^17-row
    ones at the ends), Const and Cutoff truncation."""
    psi = _vector(n, 300 + n)
    st = MPS.from_vector(psi, trunc)
    o = from_vector(n, psi, trunc, driver=driver)
    _compare(st, o, n, psi)
    if trunc is None and n <= 18:
        assert state_overlap_err(st.contract(), psi / np.linalg.norm(psi)) < 1e-12


def test_from_vector_of_a_circuit_state(gpu):
    """A state made by gates, factored back: the same spectra as the MPS
    that made it."""
    n = 12
    made = MPS.new_qubits(n, None)
    rng = np.random.default_rng(5)
    for t in range(8):
        made.apply_unis([(k, haar4(rng)) for k in range(t % 2, n - 1, 2)])
    psi = made.contract()
    st = MPS.from_vector(psi, None)
    for b in range(n - 1):
        assert_spectrum_close(st.schmidt(b), made.schmidt(b), msg=f"bond {b}")
    assert state_overlap_err(st.contract(), psi) < 1e-12
    # and it evolves as the original under a further gate
    u = haar4(rng)
    st.apply_unitary2(5, u)
    made.apply_unitary2(5, u)
    assert_spectrum_close(st.schmidt(5), made.schmidt(5))


def test_from_vector_errors_and_capacity(gpu):
    with pytest.raises(MPSError) as ei:
        MPS.from_vector(np.zeros(16))
    assert ei.value.kind == "InvalidArgument"
    with pytest.raises(MPSError) as ei:
        MPS.from_vector(np.ones(12))
    assert ei.value.kind == "StateIncompatibleShape"
    # n = 20 untruncated needs 1024 values at the middle bond: the bound is
    # 512, the state is made and the call says so
    psi = _vector(20, 1)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        st = MPS.from_vector(psi, None)
    assert any(issubclass(x.category, MPSWarning) for x in w)
    assert st.take_flags() & 2
    assert max(st.bond_dims()) == 512
    for b in range(19):
        assert abs(np.sum(st.schmidt(b) ** 2) - 1.0) < 1e-12
