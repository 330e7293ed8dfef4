"""Dense state-vector brute force used to pin the MPS oracle (test only).

Independent of the Gamma-Lambda algebra: the state is a flat complex128 vector
of 2^N amplitudes with qubit 0 as the most significant bit, which is the
ordering ``MPS::contract`` produces (mps.rs:794-798) and the basis convention
of two-qubit gates (qubit k is the MSB of the 4x4, gate.rs:750-761).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np


def product_zero(n: int) -> np.ndarray:
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def apply1(psi: np.ndarray, n: int, k: int, op: np.ndarray) -> np.ndarray:
    t = psi.reshape(2 ** k, 2, 2 ** (n - k - 1))
    return np.einsum("ts,asb->atb", op, t).reshape(-1)


def apply2(psi: np.ndarray, n: int, k: int, op: np.ndarray) -> np.ndarray:
    t = psi.reshape(2 ** k, 4, 2 ** (n - k - 2))
    return np.einsum("ts,asb->atb", op, t).reshape(-1)


def schmidt_values(psi: np.ndarray, n: int, b: int) -> np.ndarray:
    """Normalised singular values of psi reshaped (2^{b+1}, 2^{N-b-1})."""
    m = psi.reshape(2 ** (b + 1), 2 ** (n - b - 1))
    s = np.linalg.svd(m, compute_uv=False)
    return s / np.linalg.norm(s)


def entropy_from_rho(psi: np.ndarray, n: int, b: int) -> float:
    """-Tr rho_A ln rho_A from the reduced density matrix (eigh), no SVD."""
    m = psi.reshape(2 ** (b + 1), 2 ** (n - b - 1))
    m = m / np.linalg.norm(m)
    rho = m @ m.conj().T
    w = np.linalg.eigvalsh(rho)
    w = w[w > 1e-300]
    return float(-np.sum(w * np.log(w)))


def renyi_from_rho(psi: np.ndarray, n: int, b: int, a: float) -> float:
    """ln Tr rho_A^a / (1 - a) from the eigenvalues of the reduced density
    matrix (eigh), no SVD; a == 1 is the von Neumann limit.  rho_A has rank
    at most min(2^{b+1}, 2^{N-b-1}); its zero eigenvalues come out of eigh as
    round-off near 1e-17, which would dominate a < 1, so eigenvalues below
    1e-13 of the largest are taken as the exact zeros they are."""
    if a == 1.0:
        return entropy_from_rho(psi, n, b)
    m = psi.reshape(2 ** (b + 1), 2 ** (n - b - 1))
    m = m / np.linalg.norm(m)
    w = np.linalg.eigvalsh(m @ m.conj().T)
    w = w[w > 1e-13 * w.max()]
    return float(math.log(np.sum(w ** a)) / (1.0 - a))


def rho_part(psi: np.ndarray, n: int, a: int, b: int) -> np.ndarray:
    """Partial trace over every qubit outside a..b-1: rho[i, j] =
    sum_{l, r} psi[l, i, r] conj(psi[l, j, r]) with psi read as
    (2^a, 2^(b-a), 2^(n-b)); qubit a is the most significant bit of i."""
    m = psi.reshape(2 ** a, 2 ** (b - a), 2 ** (n - b))
    return np.einsum("lir,ljr->ij", m, m.conj())


def probs(psi: np.ndarray, n: int, k: int) -> Tuple[float, float]:
    t = psi.reshape(2 ** k, 2, 2 ** (n - k - 1))
    p = np.sum(np.abs(t) ** 2, axis=(0, 2))
    return float(p[0]), float(p[1])


def project(psi: np.ndarray, n: int, k: int, out: int) -> Tuple[np.ndarray, float]:
    t = psi.reshape(2 ** k, 2, 2 ** (n - k - 1)).copy()
    prob = float(np.sum(np.abs(t[:, out, :]) ** 2))
    t[:, 1 - out, :] = 0.0
    t /= math.sqrt(prob)
    return t.reshape(-1), prob


def run_dense(n: int, layers, forced: Sequence[Sequence[Tuple[int, int]]], bond: int):
    """Replay unitaries and a forced record; measurements of a layer are applied
    in increasing site order (the MPS layer's processing order, mps.rs:2218).
    Returns (probs per layer, entropy per layer incl. initial, final psi)."""
    psi = product_zero(n)
    ent = [entropy_from_rho(psi, n, bond)]
    all_probs = []
    for t, (unis, _) in enumerate(layers):
        for k, mat in unis:
            mat = np.asarray(mat)
            psi = apply2(psi, n, k, mat) if mat.shape == (4, 4) else apply1(psi, n, k, mat)
        lp = []
        for k, o in sorted(forced[t], key=lambda x: x[0]):
            psi, pr = project(psi, n, k, o)
            lp.append((k, pr))
        all_probs.append(lp)
        ent.append(entropy_from_rho(psi, n, bond))
    return all_probs, ent, psi
