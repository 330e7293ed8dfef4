"""The single-trajectory workload at its full size: N=64, Const(256), Haar
brickwork with Z measurements at p=0.1, brought to saturation on the device.

A whole trajectory at this size is hours of oracle time, so the state is
checked in two ways:
- size-independent invariants of every bond and site (normalised Schmidt
  values, p0 + p1 = 1, an entropy consistent with its spectrum);
- parity of single updates at the full size: the device state is copied
  into the oracle, and the next updates (512 x 512 thetas, 512 singular
  values truncated to 256) run on both from the same Gamma and lambda.
"""
import numpy as np
import pytest

from helpers import assert_spectrum_close
from oracle.mps_oracle import VidalMPS
from tnmps import MPS
from tnmps.workload import haar4, haar_circuit_fast

pytestmark = pytest.mark.gpu

N, CHI = 64, 256


@pytest.fixture(scope="module")
def saturated(gpu):
    c = haar_circuit_fast(N, 14, 0.1, 10546)
    st = MPS.new_qubits(N, ("const", CHI))
    st.run_circuit(c.layers, rs=iter(c.draws))
    assert st.max_bond_dim() == CHI
    return st


def _to_oracle(st):
    o = VidalMPS(N, ("const", CHI))
    o.G = [np.ascontiguousarray(st.gamma(k)) for k in range(N)]
    o.L = [st.schmidt(b) for b in range(N - 1)]
    return o


def test_fullsize_invariants(saturated):
    st = saturated
    dims = st.bond_dims()
    for b in range(N - 1):
        lam = st.schmidt(b)
        assert len(lam) == dims[b] <= CHI
        assert abs(float(np.sum(lam * lam)) - 1.0) < 1e-12, b
        assert np.all(np.diff(lam) <= 0.0), b  # descending, as LAPACK returns them
        s2 = lam * lam
        ent = -float(np.sum(s2 * np.log(s2)))
        assert abs(st.entropy_vn(b) - ent) <= 1e-12 * max(1.0, ent), b
    for k in range(0, N, 7):
        p0, p1 = st.probs(k)
        assert abs(p0 + p1 - 1.0) < 1e-12, k


def test_fullsize_updates_match_oracle(saturated):
    """Two gates on neighbouring bulk bonds from the same state: spectra at
    1e-10 relative to the leading value, probabilities and the half-chain
    entropy at 1e-10 relative."""
    st = saturated.clone()
    o = _to_oracle(st)
    d = st.bond_dims()
    # the two neighbouring bonds whose thetas are largest (measurements leave
    # some bulk bonds below chi)
    k0 = max(range(1, N - 3), key=lambda k: (d[k - 1] * d[k + 1] + d[k] * d[k + 2], -k))
    assert d[k0 - 1] == d[k0 + 1] == CHI, d
    rng = np.random.default_rng(64)
    for k in (k0, k0 + 1):
        u = haar4(rng)
        st.apply_unitary2(k, u)
        o.apply_unitary2(k, u)
        assert st.bond_dim(k) == o.bond_dim(k)
        assert_spectrum_close(st.schmidt(k), o.L[k], msg=f"bond {k}")
    assert st.bond_dim(k0) == CHI  # a 512 x 512 theta truncated to 256 values
    for k in (k0, k0 + 1, k0 + 2):
        for a, b in zip(st.probs(k), o.probs(k)):
            assert abs(a - b) <= 1e-12 + 1e-10 * abs(b), k
    e_g, e_o = st.entropy_vn(N // 2 - 1), o.entropy_vn(N // 2 - 1)
    assert abs(e_g - e_o) <= 1e-10 * abs(e_o)
