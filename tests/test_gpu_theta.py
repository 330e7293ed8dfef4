"""The two-site contraction on the matrix cores: both tile sizes, the three
contraction precisions and the launch-shape table.

The tiles are checked against the oracle (same tolerances as
tests/test_gpu_parity.py) at bond dimensions that leave partial tiles, and
against each other bitwise: a tile size changes which workgroup computes an
element of theta, not the order of its sums.
"""
import numpy as np
import pytest

from helpers import assert_spectrum_close
from oracle.mps_oracle import run_trajectory
from tnmps import MPS, Engine
from tnmps.mps import MPSError
from tnmps.workload import haar_circuit

pytestmark = pytest.mark.gpu

BUCKETS = (16, 32, 64, 128, 256, 512, 1024)


def table(tile, bw=8, split_max=8, split_rows=128, ahead=2, qr=1):
    return [dict(c_max=c, theta_tile=tile, bw=bw, split_max=split_max, split_rows=split_rows,
                 ahead=ahead, qr=qr) for c in BUCKETS]


def test_theta_tiles_match_oracle_and_each_other(gpu):
    """chi = 40: theta tiles of 32 (16 v values) and 64 (32 v values) both end
    in partial tiles.  FP64 and FP32 contractions, each tile size."""
    n, depth, p, chi = 14, 10, 0.1, 40
    c = haar_circuit(n, depth, p, 404)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    runs = {}
    for prec in ("f64", "f32"):
        for tile in (32, 64):
            st = MPS.new_qubits(n, ("const", chi)).set_precision(prec).set_tuning(table(tile))
            got = st.run_circuit(c.layers, forced=ref["record"])
            runs[prec, tile] = (st, got)
    for tile in (32, 64):
        st, got = runs["f64", tile]
        assert got["record"] == ref["record"]
        assert st.bond_dims() == ref["state"].bond_dims()
        for b in range(n - 1):
            assert_spectrum_close(st.schmidt(b), ref["state"].L[b], msg=f"tile {tile} bond {b}")
        for a, w in zip(got["entropy"], ref["entropy"]):
            assert abs(a - w) <= 1e-12 + 1e-10 * abs(w)
    for prec in ("f64", "f32"):
        a, b = runs[prec, 32][0], runs[prec, 64][0]
        assert a.bond_dims() == b.bond_dims()
        for bd in range(n - 1):
            assert np.array_equal(a.schmidt(bd), b.schmidt(bd)), f"{prec} bond {bd}"


def _prec_error(prec, seed):
    n, depth, p, chi = 12, 12, 0.1, 16
    c = haar_circuit(n, depth, p, seed)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    st = MPS.new_qubits(n, ("const", chi)).set_precision(prec)
    got = st.run_circuit(c.layers, forced=ref["record"])
    ent = max(abs(a - b) for a, b in zip(got["entropy"], ref["entropy"]))
    lam = 0.0
    for b in range(n - 1):
        x, y = st.schmidt(b), ref["state"].L[b]
        k = min(len(x), len(y))
        lam = max(lam, float(np.max(np.abs(x[:k] - y[:k]))))
    return ent, lam


def test_bf16x3_contraction_error(gpu):
    """TN_PREC_BF16X3: each FP32 operand split into two BF16 halves, three
    BF16 products per real product, FP32 accumulation; about 16 bits of each
    product are kept, so entropies and Schmidt values stay within 5e-5 of the
    FP64 oracle over 12 layers, and the error is above the FP32 path's."""
    ent, lam = _prec_error("bf16x3", 77)
    ent32, lam32 = _prec_error("f32", 77)
    print(f"bf16x3 entropy {ent:.2e} schmidt {lam:.2e}; f32 entropy {ent32:.2e} schmidt {lam32:.2e}")
    assert ent < 5e-5 and lam < 5e-5, (ent, lam)  # measured: 1.1e-5, 6.7e-6
    assert ent > 0.0 and lam > 0.0
    assert max(ent, lam) > max(ent32, lam32)


def test_tuning_table_roundtrip_and_validation(gpu):
    eng = Engine(32)
    rows = eng.get_tuning()
    assert [r["c_max"] for r in rows] == list(BUCKETS)
    new = table(64, bw=16, split_max=2, split_rows=256, ahead=3, qr=0)[:3]
    eng.set_tuning(new)
    assert eng.get_tuning() == new
    bad = [dict(new[0], theta_tile=48), dict(new[0], bw=12), dict(new[0], split_max=0),
           dict(new[0], split_max=9), dict(new[0], split_rows=32), dict(new[0], ahead=0),
           dict(new[0], ahead=4), dict(new[0], c_max=0), dict(new[0], qr=2)]
    for r in bad:
        with pytest.raises(MPSError):
            eng.set_tuning([r])
    with pytest.raises(MPSError):  # c_max not increasing
        eng.set_tuning([new[1], new[0]])
    with pytest.raises(MPSError):
        eng.set_tuning([])
    assert eng.get_tuning() == new  # a rejected table leaves the old one


def test_tuned_engine_matches_oracle(gpu):
    """Every knob of the table away from its default at once (wide Jacobi
    panels, no row split, one sweep ahead, 64-tiles): same spectra as the
    oracle, on a batch of uneven trajectories."""
    specs = [(10, 8, 0.2, 16, 1), (12, 8, 0.1, 32, 2), (8, 8, 0.3, 8, 3)]
    circuits = [haar_circuit(n, d, p, s) for (n, d, p, _, s) in specs]
    refs = [run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
            for (n, _, _, chi, _), c in zip(specs, circuits)]
    eng = Engine(32)
    eng.set_tuning(table(64, bw=16, split_max=1, split_rows=64, ahead=1))
    states = [eng.new_qubits(n, ("const", chi)) for (n, _, _, chi, _) in specs]
    for t in range(8):
        eng.batch_bilayer(states, [c.layers[t] for c in circuits],
                          forced=[r["record"][t] for r in refs])
    for st, ref, (n, _, _, _, _) in zip(states, refs, specs):
        assert st.bond_dims() == ref["state"].bond_dims()
        for b in range(n - 1):
            assert_spectrum_close(st.schmidt(b), ref["state"].L[b])


def test_autotune_small_workload(gpu):
    """The tuner times candidates on a small saturated workload and returns a
    table the engine accepts; a trajectory under that table still gives the
    oracle's spectra."""
    from tnmps.autotune import tune

    res = tune(n=10, chi=16, p=0.2, traj=2, layers=1, repeat=1, log=lambda m: None)
    rows = res["table"]
    assert [r["c_max"] for r in rows] == list(BUCKETS)
    assert len(res["trials"]) > 10 and res["tuned_s"] <= res["default_s"]
    n, depth, p, chi = 10, 8, 0.2, 16
    c = haar_circuit(n, depth, p, 5)
    ref = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    st = MPS.new_qubits(n, ("const", chi)).set_tuning(rows)
    st.run_circuit(c.layers, forced=ref["record"])
    for b in range(n - 1):
        assert_spectrum_close(st.schmidt(b), ref["state"].L[b])


def test_contraction_perf_counts(gpu):
    """tn_perf's contraction fields: the algorithmic flops of the contracted
    updates, 8 m n K + 32 m n per gate, from the task's own m, n, K."""
    from helpers import CX, H, I2

    bell = CX @ np.kron(H, I2)
    st = MPS.new_qubits(4, ("const", 8))
    st.set_timing(True)
    st.reset_perf()
    st.apply_unitary2(1, bell)  # m = n = 2, K = 1
    st.sync()
    p = st.perf()
    assert p["contract_flops"] == 8 * 2 * 2 * 1 + 32 * 2 * 2
    assert p["ms_contract"] > 0.0
    assert st.bond_dims()[1] == 2
    st.apply_unitary2(0, bell)  # m = 2, n = 4 (bond 1 has dimension 2), K = 1
    st.sync()
    p = st.perf()
    assert p["contract_flops"] == 160 + 8 * 2 * 4 * 1 + 32 * 2 * 4


def test_mixed_precision_batch_equals_solo(gpu):
    """One engine batch holding an FP64, an FP32 and a BF16-split trajectory:
    every task converges by its own precision's tolerance, so each one gives
    what it gives alone (1e-12, as for single-precision batches)."""
    n, depth, p, chi = 12, 10, 0.15, 32
    c = haar_circuit(n, depth, p, 91)
    rec = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)["record"]
    precs = ("f64", "f32", "bf16x3")
    solo = []
    for prec in precs:
        st = MPS.new_qubits(n, ("const", chi)).set_precision(prec)
        st.run_circuit(c.layers, forced=rec)
        solo.append(st)
    eng = Engine(chi)
    batch = [eng.new_qubits(n, ("const", chi)).set_precision(prec) for prec in precs]
    for t in range(depth):
        eng.batch_bilayer(batch, [c.layers[t]] * 3, forced=[rec[t]] * 3)
    for prec, a, b in zip(precs, batch, solo):
        assert a.bond_dims() == b.bond_dims(), prec
        for bd in range(n - 1):
            assert_spectrum_close(a.schmidt(bd), b.schmidt(bd), rtol=1e-12, msg=f"{prec} bond {bd}")
