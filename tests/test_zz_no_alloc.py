"""tn_apply_u2_noalloc (the allocation-free variant of tn_apply_u2) against
tn_apply_u2, through the C ABI.

Two handles hold the same state (a clone); the same gate goes through
tn_apply_u2 on one and tn_apply_u2_noalloc on the other.  Both run the same
kernels on the same inputs, so the states must agree to the parity bar
(Schmidt values 1e-12 + 1e-10 s0, probabilities 1e-12), and in fact bitwise.
Two calls in a row, the second on a smaller theta, check that the static
scratch keeps nothing from one call that the next one reads.
"""
import ctypes as C

import numpy as np
import pytest

from tnmps import MPS, MPSError
from tnmps import _lib as L
from tnmps.workload import haar4

pytestmark = pytest.mark.gpu


def _noalloc(st: MPS, k: int, u: np.ndarray) -> int:
    buf = np.ascontiguousarray(u.astype(np.complex128)).view(np.float64).reshape(-1)
    return L.load().tn_apply_u2_noalloc(st._h, k, buf.ctypes.data_as(C.POINTER(C.c_double)))


def _grow(n, chi, layers, seed):
    st = MPS.new_qubits(n, ("const", chi))
    rng = np.random.default_rng(seed)
    for t in range(layers):
        st.apply_unis([(k, haar4(rng)) for k in range(t % 2, n - 1, 2)])
    return st


def _same_state(a: MPS, b: MPS, n: int):
    assert a.bond_dims() == b.bond_dims()
    for bond in range(n - 1):
        sa, sb = a.schmidt(bond), b.schmidt(bond)
        assert np.max(np.abs(sa - sb)) <= 1e-12 + 1e-10 * sb[0]
        np.testing.assert_array_equal(sa, sb)  # same kernels, same inputs
    for k in range(n):
        np.testing.assert_allclose(a.probs(k), b.probs(k), atol=1e-12)
        np.testing.assert_array_equal(a.gamma(k), b.gamma(k))


@pytest.mark.parametrize("n,chi,layers,big", [(12, 32, 10, 5), (18, 256, 14, 8)])
def test_noalloc_equals_apply_u2(gpu, n, chi, layers, big):
    """A bulk bond (the largest theta: 512 x 512 with the row-split Jacobi at
    chi=256), then bond 0 (the smallest theta) as the second call."""
    a = _grow(n, chi, layers, 70 + n)
    b = a.clone()
    rng = np.random.default_rng(n)
    _same_state(a, b, n)
    for k in (big, 0):
        u = haar4(rng)
        a.apply_unitary2(k, u)
        assert _noalloc(b, k, u) == L.TN_OK
        _same_state(a, b, n)
    if chi == 256:
        assert max(a.bond_dims()) == 256


def test_noalloc_out_of_range_and_limit(gpu):
    """k >= n - 1 is a no-op; bonds next to the update above 256 are refused
    with TN_INVALID_ARG and the state is untouched."""
    n = 20
    st = _grow(n, 300, 16, 91)
    dims = st.bond_dims()
    assert max(dims) > 256
    before = [st.schmidt(b) for b in range(n - 1)]
    assert _noalloc(st, n - 1, np.eye(4)) == L.TN_OK
    k = next(k for k in range(1, n - 2) if dims[k - 1] > 256 or dims[k + 1] > 256)
    assert _noalloc(st, k, haar4(np.random.default_rng(3))) == L.TN_INVALID_ARG
    assert "256" in L.last_error()
    for b in range(n - 1):
        np.testing.assert_array_equal(st.schmidt(b), before[b])
    with pytest.raises(MPSError):
        st.apply_unitary2_noalloc(k, np.eye(4))
