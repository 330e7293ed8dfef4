"""Device-side workload generation against the tests' CPU restatement
(rng_ref) and the oracle: the Haar kernel's unitaries, and trajectories
whose gates, measurement masks and draws are all generated from (seed,
trajectory id)."""
import numpy as np
import pytest

import rng_ref as R
from helpers import assert_spectra_equiv
from oracle.mps_oracle import VidalMPS, apply_bilayer
from tnmps import Engine

pytestmark = pytest.mark.gpu


def test_gen_haar4_matches_restatement(gpu):
    eng = Engine(8)
    ids = [(k, t, tr) for (k, t, tr) in [(0, 0, 0), (3, 1, 7), (11, 5, 2**40), (2**31, 9, 3)]]
    got = eng.gen_haar4(10546, ids)
    for u, (c1, c2, c3) in zip(got, ids):
        want = R.haar4(10546, c1, c2, c3)
        assert np.max(np.abs(u - want)) < 1e-13
        assert np.max(np.abs(u.conj().T @ u - np.eye(4))) < 1e-14
    many = eng.gen_haar4(3, [(k, 0, 0) for k in range(2000)])
    assert np.max(np.abs(np.einsum("nji,njk->nik", many.conj(), many) - np.eye(4))) < 1e-13


@pytest.mark.parametrize("n,depth,p,chi", [(10, 8, 0.25, 8), (9, 6, 0.4, 16)])
def test_generated_trajectories_match_oracle(gpu, n, depth, p, chi):
    seed, trajs = 777, [3, 8, 21]
    eng = Engine(chi)
    states = [eng.new_qubits(n, ("const", chi)) for _ in trajs]
    refs = [VidalMPS(n, ("const", chi)) for _ in trajs]
    for t in range(depth):
        got = eng.batch_bilayer_gen(states, seed, trajs, t, [p] * len(trajs))
        for i, tr in enumerate(trajs):
            unis = [(k, R.haar4(seed, k, t, tr)) for k in range(t % 2, n - 1, 2)]
            sites = R.mask(seed, tr, t, n, p)
            ds = R.draws(seed, tr, t, len(sites))
            want = apply_bilayer(refs[i], unis, [(k, None, ds[j]) for j, k in enumerate(sites)])
            rec, prs, ent = got[i]
            assert rec == [(k, o) for (k, o, _) in want]
            for (k, pg), (_, _, pw) in zip(prs, want):
                assert abs(pg - pw) <= 1e-12 + 1e-10 * pw
            ew = refs[i].entropy_vn(n // 2 - 1)
            assert abs(ent - ew) <= 1e-12 + 1e-10 * abs(ew)
    for st, ref in zip(states, refs):
        assert_spectra_equiv([st.schmidt(b) for b in range(n - 1)], ref.L)
