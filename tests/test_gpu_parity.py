"""Parity of the HIP path (libtnmps.so through its C ABI) with the CPU oracle.

Tolerances (FP64 throughout): Schmidt values |d| <= 1e-12 + 1e-10 * s0 per
value, Born probabilities and half-chain entropy within 1e-10 relative
(1e-12 absolute floor for values near zero).  Trajectories replay the
oracle's sampled outcome record as forced (Meas::Proj) measurements, the
reference's own replay protocol (workspace/src/haar_coev_probs.rs:60-79).
"""
import math

import numpy as np
import pytest

from helpers import CX, H, SQ2, assert_spectrum_close, state_overlap_err, xrot
from oracle import dense
from oracle.mps_oracle import VidalMPS, run_trajectory
from tnmps import MPS, MPSError
from tnmps.workload import haar4, haar_circuit

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def close(a, b, rtol=RTOL, atol=1e-12):
    return abs(a - b) <= atol + rtol * abs(b)


def compare_states(gpu: MPS, ora: VidalMPS, rtol=RTOL, what=""):
    assert gpu.bond_dims() == ora.bond_dims(), f"{what} bond dims"
    for b in range(ora.n - 1):
        assert_spectrum_close(gpu.schmidt(b), ora.L[b], rtol=rtol, msg=f"{what} bond {b}")
    for k in range(ora.n):
        gp = gpu.probs(k)
        op = ora.probs(k)
        for a, c in zip(gp, op):
            assert close(a, c, rtol=rtol), f"{what} probs site {k}: {gp} vs {op}"


# ---------------------------------------------------------- closed forms
def test_product_state(gpu):
    st = MPS.new_qubits(5, ("const", 8))
    assert st.bond_dims() == [1, 1, 1, 1]
    for b in range(4):
        assert st.entropy_vn(b) == 0.0
    for k in range(5):
        assert st.probs(k) == [1.0, 0.0]
    assert st.entropy_vn(4) is None
    assert st.probs(5) is None
    assert st.prob(0, 2) is None


def test_bell_pair(gpu):
    st = MPS.new_qubits(2, None)
    st.apply_unitary1(0, H)
    st.apply_unitary2(0, CX)
    np.testing.assert_allclose(st.schmidt(0), [SQ2, SQ2], atol=1e-15)
    assert abs(st.entropy_vn(0) - math.log(2)) < 1e-14
    np.testing.assert_allclose(st.probs(1), [0.5, 0.5], atol=1e-15)


@pytest.mark.parametrize("n", [3, 6, 9])
def test_ghz(gpu, n):
    st = MPS.new_qubits(n, ("const", 4))
    st.apply_unitary1(0, H)
    for k in range(n - 1):
        st.apply_unitary2(k, CX)
    for b in range(n - 1):
        np.testing.assert_allclose(st.schmidt(b), [SQ2, SQ2], atol=1e-14)
    assert abs(st.measure_postsel_prob(n // 2, 1) - 0.5) < 1e-14
    for k in range(n):
        np.testing.assert_allclose(st.probs(k), [0.0, 1.0], atol=1e-14)
    assert st.bond_dims() == [1] * (n - 1)


def test_xrot_cnot(gpu):
    """workspace/src/rotation_testing.rs:18-27."""
    ang = 3 * math.pi / 5
    st = MPS.new_qubits(5, ("const", 2))
    st.apply_unitary1(2, xrot(ang))
    st.apply_unitary2(2, CX)
    want = [math.cos(ang / 2) ** 2, math.sin(ang / 2) ** 2]
    for k in (2, 3):
        np.testing.assert_allclose(st.probs(k), want, atol=1e-14)


# ---------------------------------------------------- oracle comparisons
def _gate_seq(n, ngates, seed):
    rng = np.random.default_rng(seed)
    seq = []
    for _ in range(ngates):
        if rng.random() < 0.2:
            q, _ = np.linalg.qr(haar4(rng)[:2, :2])
            seq.append((int(rng.integers(n)), q))
        else:
            seq.append((int(rng.integers(n - 1)), haar4(rng)))
    return seq


def _apply_both(seq, gpu, ora):
    for k, u in seq:
        if u.shape == (2, 2):
            gpu.apply_unitary1(k, u)
            ora.apply_unitary1(k, u)
        else:
            gpu.apply_unitary2(k, u)
            ora.apply_unitary2(k, u)


@pytest.mark.parametrize("n,seed", [(4, 1), (6, 2), (8, 3), (10, 4), (12, 5)])
def test_random_gates_vs_oracle(gpu, n, seed):
    seq = _gate_seq(n, 6 * n, seed)
    g = MPS.new_qubits(n, None, chi_cap=64)
    o = VidalMPS(n, None)
    _apply_both(seq, g, o)
    compare_states(g, o, what=f"n={n}")
    assert state_overlap_err(g.contract(), o.contract()) < 1e-12


def test_batched_gate_layer_equals_sequential(gpu):
    """tn_apply_unis runs a brickwork layer as one batch; the result equals
    one-by-one application."""
    n = 10
    rng = np.random.default_rng(3)
    a = MPS.new_qubits(n, ("const", 16))
    b = MPS.new_qubits(n, ("const", 16))
    for t in range(6):
        layer = [(k, haar4(rng)) for k in range(t % 2, n - 1, 2)]
        a.apply_unis(layer)
        for k, u in layer:
            b.apply_unitary2(k, u)
    for bd in range(n - 1):
        assert_spectrum_close(a.schmidt(bd), b.schmidt(bd), rtol=1e-12)


def test_probs_postsel_vs_oracle(gpu):
    n = 7
    seq = _gate_seq(n, 30, 11)
    g = MPS.new_qubits(n, None, chi_cap=16)
    o = VidalMPS(n, None)
    _apply_both(seq, g, o)
    for k, p in [(0, 1), (3, 0), (6, 1), (2, 0)]:
        pg = g.measure_postsel_prob(k, p)
        po = o.measure_postsel_prob(k, p)
        assert close(pg, po)
        compare_states(g, o, what=f"postsel {k}")


def test_measure_prob_sampling(gpu):
    """measure_prob (mps.rs:1479-1489) with a supplied uniform draw."""
    n = 6
    seq = _gate_seq(n, 24, 12)
    g = MPS.new_qubits(n, None, chi_cap=8)
    o = VidalMPS(n, None)
    _apply_both(seq, g, o)
    for k, r in [(2, 0.3), (4, 0.9), (0, 0.5)]:
        og, pg = g.measure_prob(k, r)
        oo, po = o.measure_prob(k, r)
        assert og == oo and close(pg, po)
    compare_states(g, o)


def test_meas_layer_conventions_vs_oracle(gpu):
    """Sorting, deduplication, dropped out-of-range sites and the last-site
    schedule of apply_meas_prob_layer (mps.rs:2206-2252)."""
    n = 8
    seq = _gate_seq(n, 40, 21)
    g = MPS.new_qubits(n, None, chi_cap=16)
    o = VidalMPS(n, None)
    _apply_both(seq, g, o)
    layer = [(5, 1, 0.0), (20, 0, 0.0), (1, 0, 0.0), (5, 0, 0.0), (n - 1, None, 0.4)]
    rg = g.apply_meas_prob_layer(layer)
    ro = o.apply_meas_prob_layer(layer)
    assert [(k, x) for k, x, _ in rg] == [(k, x) for k, x, _ in ro]
    for a, b in zip(rg, ro):
        assert close(a[2], b[2])
    compare_states(g, o, what="layer")
    assert g.apply_meas_prob_layer([(99, 0, 0.0)]) == []


def _forced_run(n, depth, p, chi, seed, rtol=RTOL):
    c = haar_circuit(n, depth, p, seed)
    trunc = ("const", chi)
    ref = run_trajectory(n, c.layers, trunc, rs=c.draws)
    g = MPS.new_qubits(n, trunc)
    got = g.run_circuit(c.layers, forced=ref["record"])
    assert got["record"] == ref["record"]
    for t, (a, b) in enumerate(zip(got["entropy"], ref["entropy"])):
        assert close(a, b, rtol=rtol), f"entropy layer {t}: {a} vs {b}"
    for la, lb in zip(got["probs"], ref["probs"]):
        for (ka, pa), (kb, pb) in zip(la, lb):
            assert ka == kb and close(pa, pb, rtol=rtol), f"prob site {ka}: {pa} vs {pb}"
    compare_states(g, ref["state"], rtol=rtol, what="final")
    return g, ref


def test_trajectory_config1_forced(gpu):
    """N=16, depth 16, p=0.2, chi_max=64, Haar brickwork, forced record."""
    _forced_run(16, 16, 0.2, 64, 10546)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_trajectory_truncating(gpu, seed):
    """Low p and small chi_max so that Const truncation is active."""
    _forced_run(12, 12, 0.05, 8, seed)


def test_trajectory_sampled_record(gpu):
    """Meas::Rand with the same draws gives the oracle's outcome record."""
    n, depth = 10, 10
    c = haar_circuit(n, depth, 0.25, 4242)
    ref = run_trajectory(n, c.layers, ("const", 32), rs=c.draws)
    g = MPS.new_qubits(n, ("const", 32))
    got = g.run_circuit(c.layers, rs=c.draws)
    assert got["record"] == ref["record"]
    for a, b in zip(got["entropy"], ref["entropy"]):
        assert close(a, b)


def test_cutoff_mode(gpu):
    """BondDim::Cutoff (mps.rs:262-268)."""
    n = 10
    seq = _gate_seq(n, 50, 31)
    g = MPS.new_qubits(n, ("cutoff", 1e-3), chi_cap=32)
    o = VidalMPS(n, ("cutoff", 1e-3))
    _apply_both(seq, g, o)
    compare_states(g, o, what="cutoff")


# ------------------------------------------------------------ edge cases
def test_out_of_range_conventions(gpu):
    st = MPS.new_qubits(4, ("const", 4))
    st.apply_unitary2(3, CX)       # k >= n-1: no-op (mps.rs:1335)
    st.apply_unitary1(4, H)        # k >= n: no-op (mps.rs:1301)
    assert st.bond_dims() == [1, 1, 1]
    assert st.measure_prob(4, 0.5) is None
    assert st.measure_postsel_prob(1, 2) is None
    assert st.entropy_vn(3) is None
    with pytest.raises(MPSError) as ei:
        st.apply_unitary2(0, np.eye(2))
    assert ei.value.kind == "OperatorIncompatibleShape"


def test_single_and_two_site_chains(gpu):
    one = MPS.new_qubits(1, None)
    one.apply_unitary1(0, H)
    np.testing.assert_allclose(one.probs(0), [0.5, 0.5], atol=1e-15)
    assert one.entropy_vn(0) is None
    two = MPS.new_qubits(2, None)
    two.apply_unitary1(0, H)
    two.apply_unitary2(0, CX)
    res = two.apply_meas_prob_layer([(1, 1, 0.0)])
    assert res[0][:2] == (1, 1) and abs(res[0][2] - 0.5) < 1e-14


def test_zero_probability_postselection(gpu):
    """The reference divides by 0 with no error (mps.rs:1505-1507); the ABI
    reports ZeroProbability and leaves the same NaN state."""
    st = MPS.new_qubits(3, None)
    with pytest.raises(MPSError) as ei:
        st.measure_postsel_prob(1, 1)
    assert ei.value.kind == "ZeroProbability"


def test_clone_is_independent(gpu):
    st = MPS.new_qubits(4, None)
    st.apply_unitary1(0, H)
    st.apply_unitary2(0, CX)
    cp = st.clone()
    cp.measure_postsel_prob(0, 1)
    np.testing.assert_allclose(st.probs(1), [0.5, 0.5], atol=1e-14)
    np.testing.assert_allclose(cp.probs(1), [0.0, 1.0], atol=1e-14)


def test_large_bond_single_update_row_split(gpu):
    """A lone update with theta up to 128 x 128 runs the row-split Jacobi step
    (several workgroups per block pair); spectra match the oracle."""
    n = 14
    seq = _gate_seq(n, 120, 41)
    g = MPS.new_qubits(n, ("const", 64))
    o = VidalMPS(n, ("const", 64))
    _apply_both(seq, g, o)
    compare_states(g, o, what="row split")
    for k, p in [(3, 0), (7, 1)]:
        assert close(g.measure_postsel_prob(k, p), o.measure_postsel_prob(k, p))
    compare_states(g, o, what="row split after sweeps")
