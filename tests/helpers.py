"""Shared helpers for the tests: fixed gates and tolerance checks."""
from __future__ import annotations

import numpy as np

SQ2 = 1.0 / np.sqrt(2.0)
# gate matrices as in tensor-net/lib/gate.rs:606-870 (qubit k is the MSB of
# a two-qubit operator, gate.rs:750-761)
H = np.array([[SQ2, SQ2], [SQ2, -SQ2]], dtype=np.complex128)
I2 = np.eye(2, dtype=np.complex128)
CX = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128)
CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
S = np.diag([1, 1j]).astype(np.complex128)


def xrot(theta: float) -> np.ndarray:
    """``make_xrot`` (gate.rs:718-731): e^{i theta/2} exp(-i theta X / 2)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    ph = np.exp(0.5j * theta)
    return ph * np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def kron(a, b):
    return np.kron(a, b)


def assert_spectrum_close(got, want, rtol=1e-10, atol=1e-12, msg=""):
    """Schmidt spectra: same length; |diff| <= atol + rtol * s0 elementwise.

    The tail near the truncation cutoff is only defined to round-off relative
    to the leading value (LAPACK zgesvd, and any SVD, is absolutely accurate),
    so the bound scales with the leading value s0.
    """
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{msg} length {got.shape} vs {want.shape}"
    s0 = float(want[0]) if want.size else 1.0
    err = np.max(np.abs(got - want)) if want.size else 0.0
    assert err <= atol + rtol * s0, f"{msg} spectrum error {err:.3e}"


def state_overlap_err(a: np.ndarray, b: np.ndarray) -> float:
    """1 - |<a|b>| / (|a||b|): zero iff equal up to a global phase."""
    return float(1.0 - abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b)))


def assert_spectra_equiv(got_spectra, want_spectra, rtol=1e-10, atol=1e-12, tail=1e-13):
    """Schmidt spectra per bond, equal up to round-off, where one side may
    keep extra values at the level of round-off: a singular value of an
    exactly rank-deficient theta comes out as noise near f64::EPSILON, and
    whether it passes the rank rule's EPSILON threshold (mps.rs:255-262)
    depends on the SVD's rounding, in the reference as here."""
    assert len(got_spectra) == len(want_spectra)
    for b, (g, w) in enumerate(zip(got_spectra, want_spectra)):
        g, w = np.asarray(g), np.asarray(w)
        s0 = float(max(w[0] if w.size else 0.0, g[0] if g.size else 0.0, 1e-300))
        m = min(g.size, w.size)
        assert np.all(g[m:] <= tail * s0) and np.all(w[m:] <= tail * s0), \
            f"bond {b}: lengths {g.size} vs {w.size} beyond round-off"
        if m:
            err = float(np.max(np.abs(g[:m] - w[:m])))
            assert err <= atol + rtol * s0, f"bond {b}: spectrum error {err:.3e}"
