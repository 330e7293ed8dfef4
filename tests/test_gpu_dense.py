"""Dense cross-checks on the device: the state-vector contraction
(``tn_contract``, MPS::contract mps.rs:794-798) against a state vector
evolved densely by the same gates (oracle/dense.py), and batched entropy
read-outs (``tn_entropies``) against the oracle's entropy_vn /
entropy_ry_schmidt (mps.rs:1251-1282) and the entropy of the dense
reduced density matrix."""
import math

import numpy as np
import pytest

from helpers import state_overlap_err
from oracle import dense
from oracle.mps_oracle import VidalMPS
from tnmps import Engine, MPS
from tnmps.workload import haar4

pytestmark = pytest.mark.gpu


def _brickwork(n, depth, seed):
    rng = np.random.default_rng(seed)
    return [[(k, haar4(rng)) for k in range(t % 2, n - 1, 2)] for t in range(depth)]


@pytest.mark.parametrize("n,depth", [(6, 5), (12, 8), (18, 6)])
def test_contract_matches_dense(gpu, n, depth):
    layers = _brickwork(n, depth, n)
    st = MPS.new_qubits(n, None, chi_cap=min(512, 1 << (n // 2)))
    psi = dense.product_zero(n)
    for layer in layers:
        st.apply_unis(layer)
        for k, u in layer:
            psi = dense.apply2(psi, n, k, u)
    got = st.contract()
    assert got.shape == (1 << n,)
    assert abs(np.linalg.norm(got) - 1.0) < 1e-12
    assert state_overlap_err(got, psi) < 1e-12
    # gauge-independent: the amplitudes agree up to one global phase
    i = int(np.argmax(np.abs(psi)))
    ph = got[i] / psi[i]
    assert np.max(np.abs(got - ph * psi)) < 1e-12


def test_contract_product_and_bell(gpu):
    st = MPS.new_qubits(3, ("const", 4))
    v = st.contract()
    assert v[0] == 1 and np.count_nonzero(v) == 1
    h = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    cx = np.eye(4)[[0, 1, 3, 2]]
    st.apply_unitary1(0, h)
    st.apply_unitary2(0, cx)
    v = st.contract()
    want = np.zeros(8, dtype=complex)
    want[0] = want[6] = 1 / math.sqrt(2)
    assert np.max(np.abs(v - want)) < 1e-15


def test_batched_entropies(gpu):
    n = 10
    eng = Engine(32)
    states, refs = [], []
    for s in range(3):
        st = eng.new_qubits(n, ("const", 32))
        ref = VidalMPS(n, ("const", 32))
        for layer in _brickwork(n, 4 + s, 40 + s):
            st.apply_unis(layer)
            for k, u in layer:
                ref.apply_unitary2(k, u)
        states.append(st)
        refs.append(ref)
    bonds, alphas, hs, want = [], [], [], []
    for i, ref in enumerate(refs):
        for b in (0, 3, n // 2 - 1, n - 2, n - 1, n + 5):
            for a in (1.0, 0.5, 2.0, 3.0):
                hs.append(states[i])
                bonds.append(b)
                alphas.append(a)
                w = ref.entropy_ry_schmidt(a, b)
                want.append(float("nan") if w is None else w)
    got = eng.entropies(hs, bonds, alphas)
    want = np.asarray(want)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.all(np.abs(got[ok] - want[ok]) <= 1e-12 + 1e-10 * np.abs(want[ok]))
    # von Neumann entropy equals that of the dense reduced density matrix
    psi = states[0].contract()
    for b in (0, 3, n // 2 - 1, n - 2):
        i = bonds.index(b) + alphas[bonds.index(b):].index(1.0)
        assert abs(got[i] - dense.entropy_from_rho(psi, n, b)) < 1e-10
