"""Pins of the CPU oracle (oracle/mps_oracle.py) against closed forms,
invariants and a dense state-vector brute force (oracle/dense.py).

The reference (whooie/tensor-net) has no numeric golden vectors for this path
and its Rust crate cannot be built here, so each oracle function is pinned by
mathematics chosen so that a dropped lambda weight, a wrong index or a
transposed operand fails at least one test.  The ``test_planted_*`` tests
plant such mistakes in a copy of the oracle and check that the pins catch
them.
"""
import math

import numpy as np
import pytest

from helpers import CX, CZ, H, I2, S, SQ2, assert_spectrum_close, kron, state_overlap_err, xrot
from oracle import dense
from oracle.mps_oracle import (F64_EPS, VidalMPS, from_vector, process_svd, run_trajectory,
                               truncation_rank)
from tnmps.workload import brickwork_sites, haar4, haar_circuit


# ------------------------------------------------------------ closed forms
def test_product_state():
    """MPS::new_qubits (mps.rs:370-391): lambda = [1], S = 0, probs = [1, 0]."""
    st = VidalMPS(5, ("const", 8))
    assert st.bond_dims() == [1, 1, 1, 1]
    for b in range(4):
        assert st.entropy_vn(b) == 0.0
    for k in range(5):
        assert st.probs(k) == [1.0, 0.0]
    assert st.entropy_vn(4) is None and st.probs(5) is None


def test_bell_pair():
    """H(0) CX(0,1) gives lambda[0] = (1/sqrt2, 1/sqrt2), S = ln 2."""
    st = VidalMPS(2, None)
    st.apply_unitary1(0, H)
    st.apply_unitary2(0, CX)
    np.testing.assert_allclose(st.L[0], [SQ2, SQ2], atol=1e-15)
    assert abs(st.entropy_vn(0) - math.log(2)) < 1e-14
    np.testing.assert_allclose(st.probs(0), [0.5, 0.5], atol=1e-15)


@pytest.mark.parametrize("n", [3, 6, 9])
def test_ghz_every_bond(n):
    """GHZ: every bond carries (1/sqrt2, 1/sqrt2); measuring one qubit fixes
    all others (workspace/src/testing.rs:76-82)."""
    st = VidalMPS(n, ("const", 4))
    st.apply_unitary1(0, H)
    for k in range(n - 1):
        st.apply_unitary2(k, CX)
    for b in range(n - 1):
        np.testing.assert_allclose(st.L[b], [SQ2, SQ2], atol=1e-14)
    pr = st.measure_postsel_prob(n // 2, 1)
    assert abs(pr - 0.5) < 1e-14
    for k in range(n):
        np.testing.assert_allclose(st.probs(k), [0.0, 1.0], atol=1e-14)
    assert all(d == 1 for d in st.bond_dims())


def test_xrot_cnot_probs():
    """workspace/src/rotation_testing.rs:18-27: XRot(3pi/5) on qubit 2 then
    CX(2,3) gives p0 = cos^2(3pi/10), p1 = sin^2(3pi/10) on qubits 2 and 3."""
    ang = 3 * math.pi / 5
    st = VidalMPS(5, ("const", 2))
    st.apply_unitary1(2, xrot(ang))
    st.apply_unitary2(2, CX)
    want = [math.cos(ang / 2) ** 2, math.sin(ang / 2) ** 2]
    for k in (2, 3):
        np.testing.assert_allclose(st.probs(k), want, atol=1e-14)
    for k in (0, 1, 4):
        np.testing.assert_allclose(st.probs(k), [1.0, 0.0], atol=1e-14)


def test_stabilizer_probs_closed_form():
    """Stabilizer states have outcome probabilities in {0, 1/2, 1}
    (the closed form behind workspace/src/clifford_testing.rs:172-197)."""
    rng = np.random.default_rng(7)
    gates1 = [H, S, I2]
    gates2 = [CX, CZ]
    n = 5
    st = VidalMPS(n, None)
    for _ in range(40):
        if rng.random() < 0.5:
            st.apply_unitary1(int(rng.integers(n)), gates1[rng.integers(3)])
        else:
            st.apply_unitary2(int(rng.integers(n - 1)), gates2[rng.integers(2)])
    for k in range(n):
        for p in st.probs(k):
            assert min(abs(p - v) for v in (0.0, 0.5, 1.0)) < 1e-12


# ------------------------------------------------------------- rank rule
def test_rank_rule_const_and_cutoff():
    """process_svd's rank rule (mps.rs:250-274) on synthetic spectra."""
    s = np.array([0.9, 0.4, 1e-3, 3e-16, 1e-17, 0.0])
    s = s / np.linalg.norm(s)
    assert truncation_rank(s, ("const", 10)) == 4      # 3e-16 > eps survives
    assert truncation_rank(s, ("const", 2)) == 2
    assert truncation_rank(s, ("const", 0)) == 1       # max(r, 1)
    assert truncation_rank(s, ("cutoff", 1e-2)) == 2
    assert truncation_rank(s, ("cutoff", -1e-2)) == 2  # |eps|
    assert truncation_rank(s, ("cutoff", 0.5)) == 1
    assert truncation_rank(s, None) == 4               # Cutoff(f64::EPSILON)
    # take_while stops at the first failing value, even if later ones pass
    assert truncation_rank(np.array([0.9, 0.0, 0.4]), None) == 1
    # Const counts nothing when nothing passes (no max(1) on the count)
    assert truncation_rank(np.array([1e-300 * 1e-10, 0.0]), ("const", 3)) == 0
    assert truncation_rank(np.array([1e-300 * 1e-10, 0.0]), None) == 1
    # exactly eps is not > eps
    assert truncation_rank(np.array([1.0, F64_EPS]), None) == 1


def test_process_svd_scaling():
    """q rows scale by s_j * norm2 with s_j renormalised (mps.rs:277-290)."""
    u = np.eye(3, dtype=np.complex128)
    s = np.array([3.0, 2.0, 1e-20])
    q = np.eye(3, dtype=np.complex128)
    uu, ss, qq, r = process_svd(u, s, q, ("const", 5))
    assert r == 2
    n1 = math.sqrt(9 + 4 + 1e-40)
    sn = np.array([3.0, 2.0]) / n1
    norm2 = math.sqrt(float(np.sum(sn ** 2)))
    np.testing.assert_allclose(ss, sn / norm2, rtol=1e-15)
    np.testing.assert_allclose(np.diag(qq[:, :2]).real, (sn / norm2) * norm2, rtol=1e-15)
    assert uu.shape == (3, 2)


# -------------------------------------------------------- dense brute force
def _random_gate_run(n, ngates, seed, trunc=None):
    rng = np.random.default_rng(seed)
    st = VidalMPS(n, trunc)
    psi = dense.product_zero(n)
    for _ in range(ngates):
        if rng.random() < 0.2:
            k = int(rng.integers(n))
            u = haar4(rng)[:2, :2]
            q, _ = np.linalg.qr(u)
            st.apply_unitary1(k, q)
            psi = dense.apply1(psi, n, k, q)
        else:
            k = int(rng.integers(n - 1))
            u = haar4(rng)
            st.apply_unitary2(k, u)
            psi = dense.apply2(psi, n, k, u)
    return st, psi


@pytest.mark.parametrize("n,seed", [(4, 1), (6, 2), (8, 3), (10, 4)])
def test_gates_vs_dense(n, seed):
    """Untruncated gate updates (mps.rs:1298-1387) equal the dense
    evolution: state up to phase, Schmidt values at every bond, entropy."""
    st, psi = _random_gate_run(n, 6 * n, seed)
    assert state_overlap_err(st.contract(), psi) < 1e-13
    for b in range(n - 1):
        sd = dense.schmidt_values(psi, n, b)
        sd = sd[sd > 1e-14]
        assert_spectrum_close(st.L[b], sd, rtol=1e-12, msg=f"bond {b}")
        assert abs(st.entropy_vn(b) - dense.entropy_from_rho(psi, n, b)) < 1e-12


def test_probs_and_postsel_vs_dense():
    """local_probs (mps.rs:1079-1134) and post-selection + sweep
    (mps.rs:1501-1510) equal the dense projection."""
    n = 7
    st, psi = _random_gate_run(n, 30, 11)
    for k in range(n):
        np.testing.assert_allclose(st.probs(k), dense.probs(psi, n, k), atol=1e-13)
        assert abs(sum(st.probs(k)) - 1.0) < 1e-12   # sampling_testing.rs:73
    for k, o in [(0, 1), (3, 0), (6, 1), (2, 0)]:
        pr = st.measure_postsel_prob(k, o)
        psi, prd = dense.project(psi, n, k, o)
        assert abs(pr - prd) < 1e-12
        assert state_overlap_err(st.contract(), psi) < 1e-12
        for b in range(n - 1):
            sd = dense.schmidt_values(psi, n, b)
            assert_spectrum_close(st.L[b], sd[: len(st.L[b])], rtol=1e-11, msg=f"bond {b}")


def test_refactor_sweep_is_gauge_change():
    """An untruncated refactor sweep (mps.rs:894-897) leaves observables."""
    n = 8
    st, psi = _random_gate_run(n, 30, 5)
    before = [st.L[b].copy() for b in range(n - 1)]
    pr = [st.probs(k) for k in range(n)]
    st.refactor_sweep()
    for b in range(n - 1):
        assert_spectrum_close(st.L[b], before[b], rtol=1e-12)
    for k in range(n):
        np.testing.assert_allclose(st.probs(k), pr[k], atol=1e-13)
    assert state_overlap_err(st.contract(), psi) < 1e-13


def _canonical_errors(st):
    """Left/right isometry of lambda Gamma / Gamma lambda at every site."""
    worst = 0.0
    for k in range(st.n):
        g = st.G[k]
        lL = st.L[k - 1] if k > 0 else np.ones(1)
        lR = st.L[k] if k < st.n - 1 else np.ones(1)
        a = (lL[:, None, None] * g).reshape(-1, g.shape[2])
        b = (g * lR[None, None, :]).reshape(g.shape[0], -1)
        worst = max(worst, np.abs(a.conj().T @ a - np.eye(g.shape[2])).max(),
                    np.abs(b @ b.conj().T - np.eye(g.shape[0])).max())
    return worst


@pytest.mark.parametrize("n", [6, 8, 10])
def test_trajectory_vs_dense(n):
    """Haar brickwork + measurement layers (circuit.rs:624-659,
    mps.rs:2206-2252), untruncated: entropies and Born probabilities equal the
    dense replay of the same record; the state stays canonical.

    Measurements of the last qubit are left out here: the reference's layer
    schedule does not refactor bond n-2 after projecting site n-1 (see
    test_last_site_measurement_quirk)."""
    c = haar_circuit(n, n, 0.3, 10546 + n)
    c.layers = [(u, [k for k in m if k < n - 1]) for u, m in c.layers]
    res = run_trajectory(n, c.layers, None, rs=c.draws)
    pr, ent, psi = dense.run_dense(n, c.layers, res["record"], n // 2 - 1)
    np.testing.assert_allclose(res["entropy"], ent, atol=1e-12)
    pm = [p for layer in res["probs"] for _, p in layer]
    pd = [p for layer in pr for _, p in layer]
    np.testing.assert_allclose(pm, pd, atol=1e-12)
    assert state_overlap_err(res["state"].contract(), psi) < 1e-12
    assert _canonical_errors(res["state"]) < 1e-10


def test_last_site_measurement_quirk():
    """apply_meas_prob_layer (mps.rs:2244-2250) refactors bonds last..n-2 and
    then (0..n-2).rev(): when the last processed site is n-1 the first range is
    empty and bond n-2 is never refactored.  lambda[n-2] keeps its old values
    (no longer the Schmidt values), while the contracted state still equals the
    dense projection up to phase and norm, because every refactor preserves
    the product lambda Gamma lambda Gamma lambda."""
    n = 6
    st, psi = _random_gate_run(n, 30, 8)
    lam_before = st.L[n - 2].copy()
    res = st.apply_meas_prob_layer([(n - 1, 1, 0.0)])
    psi, prd = dense.project(psi, n, n - 1, 1)
    assert abs(res[0][2] - prd) < 1e-12
    assert np.array_equal(st.L[n - 2], lam_before)
    assert state_overlap_err(st.contract(), psi) < 1e-12
    sd = dense.schmidt_values(psi, n, n - 2)
    assert np.max(np.abs(sd[: len(lam_before)] - lam_before)) > 1e-3


def test_meas_layer_sort_dedup_drop():
    """Out-of-range sites are dropped, the rest sorted and deduplicated
    keeping the first entry (mps.rs:2212-2219)."""
    n = 5
    st, _ = _random_gate_run(n, 20, 3)
    st2 = st.copy()
    res = st.apply_meas_prob_layer([(3, 1, 0.0), (9, 0, 0.0), (1, 0, 0.0), (3, 0, 0.0)])
    assert [(k, o) for k, o, _ in res] == [(1, 0), (3, 1)]
    res2 = st2.apply_meas_prob_layer([(1, 0, 0.0), (3, 1, 0.0)])
    for (a, b) in zip(res, res2):
        assert abs(a[2] - b[2]) < 1e-15
    assert st.apply_meas_prob_layer([(7, 0, 0.0)]) == []


def test_sampling_boundaries():
    """sample_state (mps.rs:1427-1447): outcome 0 iff r < p0."""
    st = VidalMPS(3, None)
    ang = 3 * math.pi / 5
    st.apply_unitary1(1, xrot(ang))
    p0 = math.cos(ang / 2) ** 2
    assert st.copy().sample_state(1, p0 * (1 - 1e-12))[0] == 0
    assert st.copy().sample_state(1, p0 * (1 + 1e-12))[0] == 1


def test_haar_and_brickwork():
    """gate::haar (gate.rs:888-909) is unitary; TileQ2 (circuit.rs:523-546)
    puts gates on bonds offs, offs+2, ... < n-1."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        u = haar4(rng)
        assert np.abs(u.conj().T @ u - np.eye(4)).max() < 1e-14
    assert brickwork_sites(6, False) == [0, 2, 4]
    assert brickwork_sites(6, True) == [1, 3]
    assert brickwork_sites(5, True) == [1, 3]


def test_truncated_trajectory_invariants():
    """Const(chi) truncation keeps sum lambda^2 = 1 and bond dims <= chi."""
    n = 12
    c = haar_circuit(n, 12, 0.1, 99)
    res = run_trajectory(n, c.layers, ("const", 8), rs=c.draws)
    st = res["state"]
    assert max(st.bond_dims()) <= 8
    for b in range(n - 1):
        assert abs(np.sum(st.L[b] ** 2) - 1.0) < 1e-13
    for k in range(n):
        assert abs(sum(st.probs(k)) - 1.0) < 1e-12


# --------------------------------------------------- Renyi entropies
RENYI_ALPHAS = (0.0, 0.5, 2.0, 3.0, 7.5)


def _renyi_two_value(c2: float, a: float) -> float:
    """S_a of the spectrum (cos t, sin t) with c2 = cos^2 t."""
    return math.log(c2 ** a + (1.0 - c2) ** a) / (1.0 - a)


@pytest.mark.parametrize("a", RENYI_ALPHAS)
def test_renyi_bell_ghz_closed_form(a):
    """entropy_ry_schmidt (mps.rs:1268-1282): a flat two-value spectrum has
    S_a = ln 2 for every a (Bell pair, and every bond of a GHZ state)."""
    st = VidalMPS(2, None)
    st.apply_unitary1(0, H)
    st.apply_unitary2(0, CX)
    assert abs(st.entropy_ry_schmidt(a, 0) - math.log(2)) < 1e-14
    g = VidalMPS(6, ("const", 4))
    g.apply_unitary1(0, H)
    for k in range(5):
        g.apply_unitary2(k, CX)
    for b in range(5):
        assert abs(g.entropy_ry_schmidt(a, b) - math.log(2)) < 1e-14
    assert g.entropy_ry_schmidt(a, 5) is None


@pytest.mark.parametrize("a", RENYI_ALPHAS)
def test_renyi_graded_two_value_closed_form(a):
    """XRot(t) then CX gives lambda = (cos t/2, sin t/2), so
    S_a = ln(cos^(2a) + sin^(2a)) / (1 - a): a graded spectrum whose value
    depends on a, and on squaring the Schmidt values before the power."""
    ang = 3 * math.pi / 5
    st = VidalMPS(4, ("const", 4))
    st.apply_unitary1(1, xrot(ang))
    st.apply_unitary2(1, CX)
    want = _renyi_two_value(math.cos(ang / 2) ** 2, a)
    assert abs(st.entropy_ry_schmidt(a, 1) - want) < 1e-13
    assert st.entropy_ry_schmidt(a, 0) == 0.0 and st.entropy_ry_schmidt(a, 2) == 0.0


@pytest.mark.parametrize("n,seed", [(6, 21), (9, 22), (12, 23)])
def test_renyi_vs_dense_rho(n, seed):
    """S_a at every bond of a random state equals ln Tr rho_A^a / (1 - a)
    from the dense reduced density matrix, and S_1 is the von Neumann
    entropy (the a == 1 branch)."""
    st, psi = _random_gate_run(n, 5 * n, seed)
    for b in range(n - 1):
        for a in RENYI_ALPHAS + (1.0,):
            want = dense.renyi_from_rho(psi, n, b, a)
            assert abs(st.entropy_ry_schmidt(a, b) - want) < 1e-11 * max(1.0, abs(want)), \
                f"bond {b} a={a}"


def test_renyi_limits():
    """S_a is non-increasing in a; S_0 = ln(rank); S_a -> S_1 as a -> 1."""
    st, _ = _random_gate_run(8, 40, 31)
    for b in range(7):
        vals = [st.entropy_ry_schmidt(a, b) for a in (0.0, 0.5, 0.999999, 2.0, 4.0)]
        assert all(x >= y - 1e-12 for x, y in zip(vals, vals[1:])), vals
        assert abs(vals[0] - math.log(st.bond_dim(b))) < 1e-12
        assert abs(vals[2] - st.entropy_vn(b)) < 1e-5


# ------------------------------------------------- from_vector / factorize
def _random_vector(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)


@pytest.mark.parametrize("n", [1, 2, 5, 9])
def test_from_vector_reproduces_the_vector(n):
    """factorize (mps.rs:448-508) untruncated: the MPS contracts back to the
    normalised vector, entry by entry, and every bond holds the dense
    Schmidt values."""
    psi = _random_vector(n, 100 + n) * 3.7
    st = from_vector(n, psi, None)
    want = psi / np.linalg.norm(psi)
    np.testing.assert_allclose(st.contract(), want, atol=1e-13)
    for b in range(n - 1):
        assert_spectrum_close(st.L[b], dense.schmidt_values(want, n, b), rtol=1e-12,
                              msg=f"bond {b}")


def test_from_vector_canonical_and_product_state():
    """The factors are left-canonical (lambda[k-1] Gamma[k] has orthonormal
    columns), and |0...0> factors into the product state of new_qubits."""
    n = 8
    st = from_vector(n, _random_vector(n, 7), None)
    for k in range(n - 1):
        lL = st.L[k - 1] if k > 0 else np.ones(1)
        a = (lL[:, None, None] * st.G[k]).reshape(-1, st.G[k].shape[2])
        assert np.abs(a.conj().T @ a - np.eye(a.shape[1])).max() < 1e-12
    z = np.zeros(2 ** 5, dtype=complex)
    z[0] = 1.0
    p = from_vector(5, z, None)
    assert p.bond_dims() == [1] * 4 and all(p.probs(k) == [1.0, 0.0] for k in range(5))


def test_from_vector_truncated():
    """Const(chi): bonds at most chi, each spectrum normalised; bond 0 is the
    dense spectrum's leading values renormalised (nothing truncated before
    it)."""
    n, chi = 10, 6
    psi = _random_vector(n, 8)
    st = from_vector(n, psi, ("const", chi))
    assert max(st.bond_dims()) <= chi
    for b in range(n - 1):
        assert abs(np.sum(st.L[b] ** 2) - 1.0) < 1e-13
    sd = dense.schmidt_values(psi / np.linalg.norm(psi), n, 0)
    np.testing.assert_allclose(st.L[0], sd[:2] / np.linalg.norm(sd[:2]), rtol=1e-12)
    with pytest.raises(Exception):
        from_vector(3, np.ones(7), None)


def test_planted_from_vector_without_division_fails():
    """Gamma[k] left undivided by the previous Schmidt values (the division of
    mps.rs:484-490 dropped) must break the contraction pin."""
    n = 6
    psi = _random_vector(n, 9)
    st = from_vector(n, psi, None)
    for k in range(1, n):
        st.G[k] = st.G[k] * st.L[k - 1][:, None, None]
    assert np.max(np.abs(st.contract() - psi / np.linalg.norm(psi))) > 1e-3


# ------------------------------------------------ into_matrix_part
@pytest.mark.parametrize("n,seed", [(6, 21), (9, 22)])
def test_matrix_part_vs_dense_partial_trace(n, seed):
    """into_matrix_part (mps.rs:1926-1943) traces out the sites outside the
    range through the canonical form only; the dense partial trace of the
    contracted vector is the independent answer, for every range with up
    to 4 sites, at the chain's ends and inside it."""
    st, psi = _random_gate_run(n, 6 * n, seed)
    for a in range(n):
        for b in range(a + 1, min(n, a + 4) + 1):
            rho = st.matrix_part(a, b)
            ref = dense.rho_part(psi, n, a, b)
            assert np.max(np.abs(rho - ref)) < 1e-12, (a, b)


def test_matrix_part_closed_forms():
    """Product |0...0>: |0><0| on every site.  Bell pair: I/2 on either
    qubit, the pure projector on both.  GHZ(5): diag(1/2, 0, ..., 0, 1/2)
    on any range of 2 or more sites.  end beyond n is clamped (mps.rs:1746)."""
    st = VidalMPS(4)
    for a in range(4):
        r = st.matrix_part(a, a + 1)
        np.testing.assert_allclose(r, np.diag([1.0, 0.0]), atol=1e-15)
    bell = VidalMPS(2)
    bell.apply_unitary1(0, H)
    bell.apply_unitary2(0, CX)
    np.testing.assert_allclose(bell.matrix_part(0, 1), I2 / 2, atol=1e-15)
    np.testing.assert_allclose(bell.matrix_part(1, 2), I2 / 2, atol=1e-15)
    v = np.array([1.0, 0.0, 0.0, 1.0]) / math.sqrt(2.0)
    np.testing.assert_allclose(bell.matrix_part(0, 2), np.outer(v, v), atol=1e-15)
    ghz = VidalMPS(5)
    ghz.apply_unitary1(0, H)
    for k in range(4):
        ghz.apply_unitary2(k, CX)
    for a, b in [(0, 2), (1, 4), (2, 5), (3, 9)]:
        d = 2 ** (min(b, 5) - a)
        want = np.zeros((d, d))
        want[0, 0] = want[-1, -1] = 0.5
        np.testing.assert_allclose(ghz.matrix_part(a, b), want, atol=1e-15)
    with pytest.raises(Exception):
        ghz.matrix_part(3, 3)


def test_matrix_part_spectrum_is_the_schmidt_spectrum():
    """For the range 0..b+1 the eigenvalues of rho are the squared Schmidt
    values of bond b (and zeros); trace 1, Hermitian."""
    n = 8
    st, _ = _random_gate_run(n, 40, 23)
    for b in range(4):
        rho = st.matrix_part(0, b + 1)
        assert np.max(np.abs(rho - rho.conj().T)) < 1e-15
        assert abs(np.trace(rho) - 1.0) < 1e-13
        w = np.sort(np.linalg.eigvalsh(rho))[::-1][: len(st.L[b])]
        np.testing.assert_allclose(w, st.L[b] ** 2, atol=1e-13)


def test_planted_matrix_part_without_boundary_weight_fails(monkeypatch):
    """The lambda^2 tracer on bond b-1 (mps.rs:1814-1826) left out must
    break the dense pin."""
    n = 7
    st, psi = _random_gate_run(n, 40, 24)
    ref = dense.rho_part(psi, n, 2, 4)
    saved = st.L[3].copy()
    st.L[3] = np.ones_like(saved)  # the tracer reads lambda[b-1] only
    bad = st.matrix_part(2, 4)
    st.L[3] = saved
    assert np.max(np.abs(bad - ref)) > 1e-3
    assert np.max(np.abs(st.matrix_part(2, 4) - ref)) < 1e-12


# ------------------------------------------- the alternate LAPACK driver
def test_gesdd_oracle_matches_gesvd_oracle():
    """The full-size GPU tests run the oracle with zgesdd for speed: on a
    truncating trajectory it gives the zgesvd oracle's records, probabilities,
    entropies and spectra to round-off."""
    n, chi = 12, 16
    c = haar_circuit(n, 12, 0.2, 4321)
    a = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    VidalMPS.svd_driver = "gesdd"
    try:
        b = run_trajectory(n, c.layers, ("const", chi), forced=a["record"])
    finally:
        VidalMPS.svd_driver = "gesvd"
    np.testing.assert_allclose(b["entropy"], a["entropy"], atol=1e-12)
    pa = [p for layer in a["probs"] for _, p in layer]
    pb = [p for layer in b["probs"] for _, p in layer]
    np.testing.assert_allclose(pb, pa, atol=1e-12)
    for bond in range(n - 1):
        assert_spectrum_close(b["state"].L[bond], a["state"].L[bond], rtol=1e-12)


# ------------------------------------------------- planted mistakes (CPU)
def test_planted_renyi_unsquared_fails(monkeypatch):
    """entropy_ry_schmidt with s^a in place of (s^2)^a (a dropped square)
    must fail the Renyi pins."""
    def bad(self, a, b):
        if a == 1.0:
            return self.entropy_vn(b)
        if b >= len(self.L):
            return None
        return math.log(sum(float(s) ** a for s in self.L[b])) / (1.0 - a)

    monkeypatch.setattr(VidalMPS, "entropy_ry_schmidt", bad)
    for a in (0.5, 2.0):
        with pytest.raises(AssertionError):
            test_renyi_graded_two_value_closed_form(a)
        with pytest.raises(AssertionError):
            test_renyi_vs_dense_rho(6, 21)


def test_planted_renyi_sign_fails(monkeypatch):
    """A wrong prefactor 1 / (a - 1) must fail the closed forms."""
    orig = VidalMPS.entropy_ry_schmidt
    monkeypatch.setattr(VidalMPS, "entropy_ry_schmidt",
                        lambda self, a, b: orig(self, a, b) if a == 1.0 else -orig(self, a, b))
    with pytest.raises(AssertionError):
        test_renyi_bell_ghz_closed_form(2.0)
def test_planted_dropped_lambda_fails(monkeypatch):
    """Dropping the lambda[k-1] row weight in contract_local2 must break the
    dense comparison."""
    orig = VidalMPS.contract_local2

    def bad(self, k):
        saved = self.L[k - 1] if k > 0 else None
        if k > 0:
            self.L[k - 1] = np.ones_like(saved)
        try:
            return orig(self, k)
        finally:
            if k > 0:
                self.L[k - 1] = saved

    monkeypatch.setattr(VidalMPS, "contract_local2", bad)
    st, psi = _random_gate_run(6, 30, 2)
    assert state_overlap_err(st.contract(), psi) > 1e-6


def test_planted_transposed_gate_fails(monkeypatch):
    """Applying U^T instead of U in apply_unitary2 must break the pins."""
    orig = VidalMPS.apply_unitary2
    monkeypatch.setattr(VidalMPS, "apply_unitary2", lambda self, k, op: orig(self, k, np.asarray(op).T))
    st, psi = _random_gate_run(6, 30, 2)
    assert state_overlap_err(st.contract(), psi) > 1e-6
