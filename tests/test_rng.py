"""The workload generator's host side (``tn_rng_words`` through the C ABI)
and the tests' restatement of it, on the CPU.

Philox4x64-10 is pinned by the Random123 known-answer vector for a zero
counter and key and by numpy's independent Philox bit generator (which
steps its counter before each block).  The Haar construction is pinned by
unitarity and by moments of the Haar measure on U(4): E|U_ij|^2 = 1/4,
E|U_ij|^4 = 2/(4*5) = 1/10, E U_ij = 0 (the last fails without the phase
correction of gate.rs:900-906)."""
import numpy as np
import pytest

import rng_ref as R
from tnmps.mps import rng_words

KAT_ZERO = [0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]


def test_philox_known_answer():
    assert R.philox((0, 0, 0, 0), (0, 0)) == KAT_ZERO
    assert [int(x) for x in rng_words(0, 0, 0, 0, 1)] == KAT_ZERO


@pytest.mark.parametrize("ctr,key", [((0, 0, 0, 0), (1, 0)), ((5, 7, 11, 13), (12345, 0)),
                                     ((R.MASK64, 3, 0, 9), (2, 0)),
                                     ((R.MASK64, R.MASK64, 4, 1), (77, 0))])
def test_philox_matches_numpy(ctr, key):
    bg = np.random.Philox(counter=np.array(ctr, dtype=np.uint64),
                          key=np.array(key, dtype=np.uint64))
    raw = [int(x) for x in bg.random_raw(8)]

    def inc(c):
        c = list(c)
        for i in range(4):
            c[i] = (c[i] + 1) & R.MASK64
            if c[i]:
                break
        return c

    c1 = inc(ctr)
    c2 = inc(c1)
    assert raw == R.philox(c1, key) + R.philox(c2, key)
    if c1[0] == 0 and c2[0] == 1 and c1[1:] == c2[1:]:
        lib = [int(x) for x in rng_words(key[0], c1[1], c1[2], c1[3], 2)]
        assert lib == raw


def test_library_words_match_restatement():
    for (seed, c1, c2, c3) in [(10546, 3, 7, 11), (1, R.MASK_STREAM, 0, 5), (2**63, R.DRAW_STREAM, 9, 2**40)]:
        assert [int(x) for x in rng_words(seed, c1, c2, c3, 5)] == R.words(seed, c1, c2, c3, 5)


def test_u01_range():
    assert R.u01(0) == 0.0
    assert R.u01(R.MASK64) == 1.0 - 2.0 ** -53


def test_haar_moments():
    us = np.stack([R.haar4(7, k, 0, 0) for k in range(4000)])
    assert np.max(np.abs(np.einsum("nji,njk->nik", us.conj(), us) - np.eye(4))) < 1e-13
    a2 = np.abs(us) ** 2
    assert abs(a2.mean() - 0.25) < 2e-3
    assert abs((a2 ** 2).mean() - 0.1) < 4e-3
    assert np.max(np.abs(us.mean(axis=0))) < 0.03
    # a Q without the phase correction is biased
    z = np.stack([np.linalg.qr(R.haar4(7, k, 0, 0) @ np.diag([1, 1j, -1, -1j]))[0] for k in range(400)])
    assert abs(z[:, 0, 0].mean()) > 0.1


def test_mask_and_draws():
    n, p = 40, 0.3
    m = R.mask(5, 3, 2, n, p)
    assert m == sorted(set(m)) and all(0 <= q < n for q in m)
    tot = sum(len(R.mask(5, t, 0, n, p)) for t in range(200))
    assert abs(tot / (200 * n) - p) < 0.02
    d = R.draws(5, 3, 2, 9)
    assert len(d) == 9 and all(0 <= x < 1 for x in d)
