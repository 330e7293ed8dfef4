"""The single-trajectory workload at its full size: N=64, Const(256), Haar
brickwork with Z measurements at p=0.1, brought to saturation on the device.

A whole trajectory at this size is hours of oracle time, so the state is
checked in three ways:
- size-independent invariants of every bond and site (normalised Schmidt
  values, p0 + p1 = 1, an entropy consistent with its spectrum);
- parity of single updates at the full size: the device state is copied
  into the oracle, and the next updates (512 x 512 thetas, 512 singular
  values truncated to 256) run on both from the same Gamma and lambda;
- parity of one whole measurement layer from the saturated state: a
  mid-chain site measured, then the L->R and R->L refactor sweeps of
  apply_meas_prob_layer (mps.rs:2206-2252), about 90 refactors that the
  device runs as factored updates (mode 1, mode 2, gauge prepasses) with the
  row-split Jacobi on 512 x 256 SVDs.

The oracle runs these with LAPACK's zgesdd (VidalMPS.svd_driver), the same
SVD as zgesvd to round-off and about 15 times faster at this size
(tests/test_oracle_pins.py::test_gesdd_oracle_matches_gesvd_oracle).
"""
import numpy as np
import pytest

from helpers import assert_spectra_equiv, assert_spectrum_close
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
    o.svd_driver = "gesdd"
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


def test_fullsize_measurement_layer_matches_oracle(saturated):
    """One forced measurement of site N/2 from the saturated state, with the
    layer's whole refactor schedule, on the device and on the oracle.
    Spectra at every bond to 1e-10 of the leading value (a value at
    round-off level that one side keeps and the other drops is allowed, as
    helpers.assert_spectra_equiv documents), probabilities at every site and
    the half-chain entropy to 1e-10 relative."""
    st = saturated.clone()
    o = _to_oracle(st)
    k = N // 2
    p = st.probs(k)
    out = 0 if p[0] >= p[1] else 1
    st.set_timing(True)
    st.reset_perf()
    res = st.apply_meas_prob_layer([(k, out, 0.0)])
    perf = st.perf()
    st.set_timing(False)
    ores = o.apply_meas_prob_layer([(k, out, 0.0)])
    assert [(a, b) for a, b, _ in res] == [(a, b) for a, b, _ in ores]
    assert abs(res[0][2] - ores[0][2]) <= 1e-10 * ores[0][2]
    # the schedule: refactors k .. N-2, then N-3 .. 0 (mps.rs:2246-2250)
    assert perf["updates"] == (N - 1 - k) + (N - 2)
    assert perf["factored"] > (N - 2)  # factored SVDs (prepasses included) ran
    assert_spectra_equiv([st.schmidt(b) for b in range(N - 1)], o.L, rtol=1e-10)
    for j in range(N):
        for a, b in zip(st.probs(j), o.probs(j)):
            assert abs(a - b) <= 1e-12 + 1e-10 * abs(b), j
    e_g, e_o = st.entropy_vn(N // 2 - 1), o.entropy_vn(N // 2 - 1)
    assert abs(e_g - e_o) <= 1e-10 * abs(e_o)
    assert st.probs(k)[out] == pytest.approx(1.0, abs=1e-12)


def test_fullsize_fp32_error_per_layer(saturated):
    """The FP32 contraction against the FP64 path on the same outcome record,
    four full layers (gates and measurements, p=0.1) from the saturated
    state: the half-chain entropy and every Schmidt value stay within 1e-6
    of the FP64 values after every layer (the error grows with depth: see
    the mixed-precision measurements of bench.py)."""
    c = haar_circuit_fast(N, 4, 0.1, 777)
    f64 = saturated.clone()
    f32 = saturated.clone().set_precision("f32")
    ref = f64.run_circuit(c.layers, rs=iter(c.draws))
    worst_e, worst_l = 0.0, 0.0
    for t in range(4):
        got = f32.run_circuit(c.layers[t:t + 1], forced=ref["record"][t:t + 1])
        worst_e = max(worst_e, abs(got["entropy"][1] - ref["entropy"][t + 1]))
    for b in range(N - 1):
        a, g = f64.schmidt(b), f32.schmidt(b)
        m = min(len(a), len(g))
        worst_l = max(worst_l, float(np.max(np.abs(a[:m] - g[:m]))))
    print(f"FP32 vs FP64, 4 layers at N=64 chi=256: entropy {worst_e:.3e}, schmidt {worst_l:.3e}")
    assert 0.0 < worst_e <= 1e-6 and worst_l <= 1e-6
