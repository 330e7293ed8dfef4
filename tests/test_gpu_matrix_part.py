"""MPS::into_matrix_part (mps.rs:1926-1943) on the device (tn_matrix_part)
against the oracle's restatement run on the device state's own Gamma and
lambda (to 1e-12 in every entry), and against the dense partial trace of
the contracted vector for an untruncated state."""
import numpy as np
import pytest

from oracle import dense
from oracle.mps_oracle import VidalMPS
from tnmps import MPS, MPSError
from tnmps.workload import haar4, haar_circuit

pytestmark = pytest.mark.gpu


def _to_oracle(st: MPS) -> VidalMPS:
    o = VidalMPS(st.n)
    o.G = [np.ascontiguousarray(st.gamma(k)) for k in range(st.n)]
    o.L = [st.schmidt(b) for b in range(st.n - 1)]
    return o


def _ranges(n, width):
    out = []
    for a in range(n):
        for w in width:
            out.append((a, a + w))
    return out


def test_matrix_part_untruncated_vs_dense(gpu):
    n = 10
    st = MPS.new_qubits(n, None)
    rng = np.random.default_rng(71)
    for t in range(6):
        st.apply_unis([(k, haar4(rng)) for k in range(t % 2, n - 1, 2)])
    psi = st.contract()
    o = _to_oracle(st)
    for a, b in _ranges(n, (1, 2, 3, 5)):
        idx, rho = st.matrix_part(a, b)
        bb = min(b, n)
        assert idx == [("Ket", k) for k in range(a, bb)] + [("Bra", k) for k in range(a, bb)]
        assert np.max(np.abs(rho - o.matrix_part(a, b))) < 1e-12, (a, b)
        assert np.max(np.abs(rho - dense.rho_part(psi, n, a, bb))) < 1e-12, (a, b)
    _, whole = st.matrix_part(0, n)
    assert np.max(np.abs(whole - np.outer(psi, psi.conj()))) < 1e-12


def test_matrix_part_truncated_trajectory(gpu):
    """N=32, Const(64) after 16 layers of gates and measurements (bulk bonds
    at 64): ranges of 1 to 8 sites, at both ends and in the bulk, against the
    oracle on the same Gamma and lambda; Hermitian with unit trace."""
    n, chi = 32, 64
    c = haar_circuit(n, 16, 0.1, 7100)
    st = MPS.new_qubits(n, ("const", chi))
    st.run_circuit(c.layers, rs=iter(c.draws))
    assert max(st.bond_dims()) == chi
    o = _to_oracle(st)
    for a, b in [(0, 1), (0, 4), (5, 6), (12, 20), (14, 16), (28, 32), (31, 40), (16, 23)]:
        _, rho = st.matrix_part(a, b)
        ref = o.matrix_part(a, b)
        assert np.max(np.abs(rho - ref)) < 1e-12, (a, b)
        assert np.max(np.abs(rho - rho.conj().T)) < 1e-14
        assert abs(np.trace(rho).real - 1.0) < 1e-10


def test_matrix_part_errors(gpu):
    st = MPS.new_qubits(16, None)
    for a, b in [(3, 3), (5, 2), (16, 20), (0, 13)]:
        with pytest.raises(MPSError) as ei:
            st.matrix_part(a, b)
        assert ei.value.kind == "InvalidArgument"
    _, rho = st.matrix_part(4, 16)  # 12 sites: the largest range
    assert rho.shape == (4096, 4096) and rho[0, 0] == 1.0 and np.count_nonzero(rho) == 1
