"""The algebra of the library's factored updates, checked on the CPU against
the oracle's full-theta refactor (mps.rs:846-880).

theta = A M with A = lamL Gamma[k] (m x K), M = lamM Gamma[k+1] lamR
(K x n).  Mode 1 (A has orthonormal columns) decomposes M; mode 2 (Gamma[k+1]
lamR has orthonormal rows) decomposes lamL Gamma[k] lamM; the gauge prepass
applies mode 2 to a factor pair where neither is an isometry and then the
real refactor runs in mode 1.  These numpy restatements use LAPACK for the
small SVDs; the point is that the factored split (which singular vectors
multiply which Gamma, which lambda divisions remain) equals the reference's.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle.mps_oracle import VidalMPS, process_svd


def _rand_state(n, chi, seed, layers=6):
    rng = np.random.default_rng(seed)
    st = VidalMPS(n, ("const", chi))
    for t in range(layers):
        for k in range(t % 2, n - 1, 2):
            g = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
            q, _ = np.linalg.qr(g)
            st.apply_unitary2(k, q)
    return st


def _lam(st, b):
    return st.L[b] if 0 <= b < st.n - 1 else None


def _factored(st, k, mode):
    """One factored update of bond k in place, as the kernels compute it."""
    G0, G1 = st.G[k], st.G[k + 1]
    chiL, _, K = G0.shape
    chiR = G1.shape[2]
    m, n = 2 * chiL, 2 * chiR
    lamL, lamM, lamR = _lam(st, k - 1), st.L[k], _lam(st, k + 1)
    if mode == 1:
        Mm = G1 * lamM[:, None, None]
        if lamR is not None:
            Mm = Mm * lamR[None, None, :]
        uM, s, vh = sla.svd(Mm.reshape(K, n), full_matrices=False, lapack_driver="gesvd")
        u = G0.reshape(m, K) @ uM                   # Gamma[k] W: lamL already divided
        u2, s2, q, rank = process_svd(u, s, vh, st.trunc)
        gk = u2.reshape(chiL, 2, rank)
        g1 = q.reshape(rank, 2, chiR) / s2[:, None, None]
        if lamR is not None:
            g1 = g1 / lamR[None, None, :]
    else:
        Ap = G0 * lamM[None, None, :]
        if lamL is not None:
            Ap = Ap * lamL[:, None, None]
        uA, s, vhA = sla.svd(Ap.reshape(m, K), full_matrices=False, lapack_driver="gesvd")
        vh = vhA @ G1.reshape(K, n)                 # W^H Gamma[k+1]: lamR already divided
        u2, s2, q, rank = process_svd(uA, s, vh, st.trunc)
        gk = u2.reshape(chiL, 2, rank)
        if lamL is not None:
            gk = gk / lamL[:, None, None]
        g1 = q.reshape(rank, 2, chiR) / s2[:, None, None]
    st.G[k] = np.ascontiguousarray(gk)
    st.L[k] = s2
    st.G[k + 1] = np.ascontiguousarray(g1)
    st.local_renormalize(k)
    st.local_renormalize(k + 1)


def _spectra_close(a, b):
    assert len(a.L) == len(b.L)
    for la, lb in zip(a.L, b.L):
        assert la.shape == lb.shape
        assert np.max(np.abs(la - lb)) <= 1e-12 + 1e-10 * lb[0]


def _state_close(a, b):
    va, vb = a.contract(), b.contract()
    ov = abs(np.vdot(va, vb)) / (np.linalg.norm(va) * np.linalg.norm(vb))
    assert 1.0 - ov < 1e-12


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mode1_after_left_factor_update(seed):
    """After any update of bond k, site k is left-fresh: refactoring bond k
    from M alone equals the full refactor."""
    st = _rand_state(8, 6, seed)
    k = 3
    st.local_refactor(k)          # makes lamL Gamma[k] an isometry
    ref, got = st.copy(), st.copy()
    ref.local_refactor(k)
    _factored(got, k, 1)
    _spectra_close(got, ref)
    _state_close(got, ref)


@pytest.mark.parametrize("seed", [4, 5])
def test_mode2_after_right_factor_update(seed):
    """After an update of bond k, site k+1 is right-fresh: a refactor of bond
    k from lamL Gamma[k] lamM alone equals the full refactor."""
    st = _rand_state(8, 6, seed)
    k = 2
    st.local_refactor(k)          # makes Gamma[k+1] lamR co-isometric
    ref, got = st.copy(), st.copy()
    ref.local_refactor(k)
    _factored(got, k, 2)
    _spectra_close(got, ref)
    _state_close(got, ref)


@pytest.mark.parametrize("seed", [6, 7, 8])
def test_gauge_prepass_then_mode1(seed):
    """Neither site fresh (a projection on each side of the bond): mode 2 as a
    gauge change, then mode 1, equals the full refactor of the original."""
    st = _rand_state(10, 8, seed)
    k = 4
    for j in (k, k + 1):          # break both isometries
        st.project(j, 0, complex(np.sqrt(max(st.local_prob(j, 0), 1e-300))))
    ref, got = st.copy(), st.copy()
    ref.local_refactor(k)
    _factored(got, k, 2)          # gauge prepass: site k becomes left-fresh
    _factored(got, k, 1)
    _spectra_close(got, ref)
    _state_close(got, ref)


def test_mode1_without_isometry_is_wrong():
    """Planted mistake: mode 1 when site k is NOT left-fresh gives a wrong
    spectrum (the freshness bookkeeping matters)."""
    st = _rand_state(10, 8, 9)
    k = 4
    st.project(k, 0, complex(np.sqrt(st.local_prob(k, 0))))
    ref, got = st.copy(), st.copy()
    ref.local_refactor(k)
    _factored(got, k, 1)
    err = max(float(np.max(np.abs(a - b))) for a, b in zip(got.L, ref.L) if a.shape == b.shape)
    assert err > 1e-6
