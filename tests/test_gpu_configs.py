"""The batched, large-bond and serving workloads at their own sizes.

- 1024 trajectories of N=32, Const(128), measurement rates on the grid
  linspace(0.05, 0.30, 11), advanced in lock step on one engine: invariants
  of every trajectory, two trajectories equal to their solo runs, and single
  updates from the most saturated one against the oracle.
- N=48, Const(512): gate updates on 1024 x 1024 thetas (the widest SVD the
  library poses) against the oracle in FP64, and the FP32 contraction's
  error on the same updates.
- Continuous batching of FP32 and FP64 requests with N in {16, 32, 64} and
  chi_max in {64, 128, 256}: every request equals its solo run (FP64 to
  1e-12, FP32 within the FP32 error bound), and the N=16, chi=64 FP32 ones stay
  within twice the FP32 bound of the FP64 oracle replaying their record over
  their 16 layers.

Oracles at 1024 x 1024 use LAPACK's zgesdd (VidalMPS.svd_driver); it is
checked against the zgesvd oracle in tests/test_oracle_pins.py.
"""
import numpy as np
import pytest

from helpers import assert_spectrum_close
from oracle.mps_oracle import VidalMPS, run_trajectory
from tnmps import MPS, Engine
from tnmps.serve import Request, Server
from tnmps.workload import haar4, haar_circuit, haar_circuit_fast

pytestmark = pytest.mark.gpu

FP32_BOUND = 1e-6  # entropy and Schmidt-value error of the FP32 contraction


def _to_oracle(st: MPS, trunc, driver="gesvd") -> VidalMPS:
    n = st.n
    o = VidalMPS(n, trunc)
    o.svd_driver = driver
    o.G = [np.ascontiguousarray(st.gamma(k)) for k in range(n)]
    o.L = [st.schmidt(b) for b in range(n - 1)]
    return o


def _bulk_pair(d, chi):
    """A bond k whose neighbours k-1 and k+1 are both at chi (the largest
    theta), with k+1 also such a bond when possible."""
    ks = [k for k in range(1, len(d) - 1) if d[k - 1] == d[k + 1] == chi]
    assert ks, d
    for k in ks:
        if k + 1 in ks:
            return k
    return ks[0]


def _check_invariants(st: MPS, chi: int, ent: float):
    n = st.n
    for b in range(n - 1):
        lam = st.schmidt(b)
        assert 1 <= len(lam) <= chi
        assert abs(float(np.sum(lam * lam)) - 1.0) < 1e-12
        assert np.all(np.diff(lam) <= 0.0)
    lam = st.schmidt(n // 2 - 1)
    s2 = lam * lam
    assert abs(ent + float(np.sum(s2 * np.log(s2)))) <= 1e-12 * max(1.0, ent)
    for k in (0, n // 2, n - 1):
        assert abs(sum(st.probs(k)) - 1.0) < 1e-12


def test_batched_1024_trajectories_pgrid(gpu):
    n, chi, T, depth = 32, 128, 1024, 10
    pgrid = np.linspace(0.05, 0.30, 11)
    circs = [haar_circuit_fast(n, depth, float(pgrid[i % 11]), 9000 + i) for i in range(T)]
    eng = Engine(chi)
    states = [eng.new_qubits(n, ("const", chi)) for _ in range(T)]
    draws = [iter(c.draws) for c in circs]
    recs = [[] for _ in range(T)]
    ents = [[] for _ in range(T)]
    for t in range(depth):
        res = eng.batch_bilayer(states, [c.layers[t] for c in circs], draws=draws)
        for i, (rec, _, e) in enumerate(res):
            recs[i].append(rec)
            ents[i].append(e)
    dims = [st.bond_dims() for st in states]
    assert max(max(d) for d in dims) == chi
    assert len(set(tuple(d) for d in dims)) > 100  # uneven bond dimensions
    for i, st in enumerate(states):
        _check_invariants(st, chi, ents[i][-1])
    # two trajectories (lowest and highest p) alone on their own engines
    for i in (0, 10):
        solo = MPS.new_qubits(n, ("const", chi))
        got = solo.run_circuit(circs[i].layers, rs=iter(circs[i].draws))
        assert got["record"] == recs[i]
        for a, b in zip(got["entropy"][1:], ents[i]):
            assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), i
        for b in range(n - 1):
            assert_spectrum_close(solo.schmidt(b), states[i].schmidt(b), rtol=1e-12)
    # single updates from the most saturated trajectory against the oracle
    j = max(range(T), key=lambda i: sum(dims[i]))
    st = states[j].clone()
    o = _to_oracle(st, ("const", chi))
    k0 = _bulk_pair(dims[j], chi)
    rng = np.random.default_rng(32)
    for k in (k0, k0 + 1):
        u = haar4(rng)
        st.apply_unitary2(k, u)
        o.apply_unitary2(k, u)
        assert st.bond_dim(k) == o.bond_dim(k)
        assert_spectrum_close(st.schmidt(k), o.L[k], msg=f"bond {k}")
    for k in (k0, k0 + 1, k0 + 2):
        for a, b in zip(st.probs(k), o.probs(k)):
            assert abs(a - b) <= 1e-12 + 1e-10 * abs(b)


def test_chi512_gate_updates_vs_oracle(gpu):
    """N=48, Const(512), brought to a bulk bond dimension of 512 by
    brickwork gate layers: two gate updates on neighbouring bulk bonds
    (1024 x 1024 thetas, 1024 singular values truncated to 512) against the
    oracle in FP64; the FP32 contraction's error on the same updates stays
    within FP32_BOUND."""
    n, chi = 48, 512
    st = MPS.new_qubits(n, ("const", chi))
    rng = np.random.default_rng(4848)
    for t in range(18):
        st.apply_unis([(k, haar4(rng)) for k in range(t % 2, n - 1, 2)])
    d = st.bond_dims()
    k0 = _bulk_pair(d, chi)
    o = _to_oracle(st, ("const", chi), driver="gesdd")
    f32 = st.clone().set_precision("f32")
    for k in (k0, k0 + 1):
        u = haar4(rng)
        st.apply_unitary2(k, u)
        o.apply_unitary2(k, u)
        f32.apply_unitary2(k, u)
        assert st.bond_dim(k) == o.bond_dim(k)
        assert_spectrum_close(st.schmidt(k), o.L[k], msg=f"bond {k}")
    assert st.bond_dim(k0) == chi
    for k in (k0, k0 + 1, k0 + 2):
        for a, b in zip(st.probs(k), o.probs(k)):
            assert abs(a - b) <= 1e-12 + 1e-10 * abs(b)
    for b in (k0, k0 + 1, n // 2 - 1):
        assert abs(st.entropy_vn(b) - o.entropy_vn(b)) <= 1e-10 * o.entropy_vn(b)
    err_e = max(abs(f32.entropy_vn(b) - o.entropy_vn(b)) for b in (k0, k0 + 1))
    err_l = 0.0
    for b in (k0, k0 + 1):
        x, y = f32.schmidt(b), o.L[b]
        m = min(len(x), len(y))
        err_l = max(err_l, float(np.max(np.abs(x[:m] - y[:m]))))
    print(f"FP32 contraction at chi=512: entropy error {err_e:.3e}, Schmidt {err_l:.3e}")
    assert 0.0 < err_e <= FP32_BOUND and err_l <= FP32_BOUND


def test_serving_mixed_precision_mix_large(gpu):
    """Requests of N in {16, 32, 64} and chi_max in {64, 128, 256}, FP32 and
    FP64 contractions, served by two lock-step workers.

    - FP64 requests equal their solo runs to 1e-12.
    - FP32 requests equal their solo runs within the FP32 error bound
      (2 FP32_BOUND over 16 layers): the FP32 contraction rounds FP64
      operands, and the FP64 Jacobi sums its Gram matrices in an order set by
      the launch shape (one workgroup per block pair in a batch, row slices
      when alone), so a round-off difference upstream can move an operand to
      the neighbouring FP32 value, and the trajectory carries that difference
      on as it carries the FP32 rounding itself (measured: 2.7e-7 in the
      entropy after 16 layers at N=32, chi=128).
    - The N=16, chi=64 FP32 requests are within 2 FP32_BOUND of the FP64
      oracle replaying their outcome record over their 16 layers.
    Depth 16 for every request keeps the test short (the serving workload
    uses depth N); the N=64 request reaches chi=256."""
    specs = [(16, 0.2, 64, 0.0, "f32"), (32, 0.1, 128, 0.1, "f32"), (64, 0.05, 256, 0.2, "f32"),
             (16, 0.35, 64, 0.3, "f32"), (32, 0.3, 256, 0.4, "f32"), (32, 0.15, 128, 0.5, "f64"),
             (16, 0.25, 64, 0.6, "f64")]
    reqs = []
    for i, (n, p, chi, t, prec) in enumerate(specs):
        c = haar_circuit(n, 16, p, 6100 + i)
        reqs.append(Request(rid=i, n=n, p=p, chi=chi, layers=c.layers, draws=c.draws,
                            arrival=t, prec=prec))
    srv = Server(chi_max=256, workers=2)
    srv.run(reqs)
    fp32_solo = 0.0
    for r in reqs:
        assert r.done is not None and len(r.record) == r.depth
        tol = 1e-12 if r.prec == "f64" else 2 * FP32_BOUND
        solo = MPS.new_qubits(r.n, ("const", r.chi)).set_precision(r.prec)
        got = solo.run_circuit(r.layers, rs=iter(r.draws))
        assert got["record"] == r.record
        for a, b in zip(got["entropy"], r.entropy):
            assert abs(a - b) <= tol * max(1.0, abs(b)), (r.rid, a, b)
            if r.prec == "f32":
                fp32_solo = max(fp32_solo, abs(a - b))
        for la, lb in zip(got["probs"], r.probs):
            for (ka, pa), (kb, pb) in zip(la, lb):
                assert ka == kb and abs(pa - pb) <= tol
        if r.n == 16 and r.prec == "f32":
            ref = run_trajectory(r.n, r.layers, ("const", r.chi), forced=r.record)
            err = max(abs(a - b) for a, b in zip(ref["entropy"], r.entropy))
            perr = max([abs(a - b) for la, lb in zip(ref["probs"], r.probs)
                        for (_, a), (_, b) in zip(la, lb)] or [0.0])
            print(f"request {r.rid}: FP32 vs FP64 oracle entropy {err:.3e}, probs {perr:.3e}")
            # the FP32 error grows with depth, about 1e-7 per layer
            assert err <= 2 * FP32_BOUND and perr <= 2 * FP32_BOUND
    print(f"FP32 requests against their solo runs: entropy {fp32_solo:.3e}")
    big = next(r for r in reqs if r.n == 64)
    assert max(big.entropy) > 1.0
