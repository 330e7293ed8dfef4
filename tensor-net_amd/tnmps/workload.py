"""Synthetic workload generation for the MPS update path.

Restates the reference's circuit generators with a seeded numpy generator
(the reference's rand_chacha/statrs streams are not reproduced; every
comparison passes explicit unitaries and a forced record instead):

* ``haar4``  -- ``gate::haar(2, rng)`` (tensor-net/lib/gate.rs:888-909): a 4x4
  complex Ginibre matrix (Re drawn before Im, row-major), Householder QR,
  column j divided by r_jj/|r_jj|.
* ``brickwork_sites`` -- ``TileQ2::new(n, offs)`` (circuit.rs:523-530): bonds
  offs, offs+2, ... < n-1; layer t is offset when t is odd
  (workspace/lib/haar.rs:285).
* ``uniform_meas`` -- circuit.rs:580-584: each qubit in order measured with
  probability p.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np


def haar4(rng: np.random.Generator) -> np.ndarray:
    x = rng.standard_normal(32).reshape(4, 4, 2)
    z = x[..., 0] + 1j * x[..., 1]
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q / (d / np.abs(d))[None, :]


def haar4_batch(rng: np.random.Generator, count: int) -> np.ndarray:
    """``count`` Haar 4x4 unitaries at once (same construction as ``haar4``,
    batched QR): shape (count, 4, 4)."""
    x = rng.standard_normal(count * 32).reshape(count, 4, 4, 2)
    z = x[..., 0] + 1j * x[..., 1]
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    return q / (d / np.abs(d))[:, None, :]


def brickwork_sites(n: int, offs: bool) -> List[int]:
    return list(range(int(offs), n - 1, 2))


def uniform_meas(n: int, p: float, rng: np.random.Generator) -> List[int]:
    return [k for k in range(n) if rng.random() < p]


@dataclass
class Circuit:
    """A brickwork Haar circuit with Bernoulli(p) measurement locations."""
    n: int
    depth: int
    p: float
    seed: int
    layers: List[Tuple[List[Tuple[int, np.ndarray]], List[int]]] = field(default_factory=list)
    draws: List[float] = field(default_factory=list)


def haar_circuit(n: int, depth: int, p: float, seed: int) -> Circuit:
    """Unitaries from stream ``seed``, measurement locations from ``seed+1``,
    uniform draws for Born sampling from ``seed+2`` (one per possible
    measurement, consumed in processing order)."""
    ru = np.random.default_rng(seed)
    rm = np.random.default_rng(seed + 1)
    rd = np.random.default_rng(seed + 2)
    c = Circuit(n=n, depth=depth, p=p, seed=seed)
    for t in range(depth):
        unis = [(k, haar4(ru)) for k in brickwork_sites(n, t % 2 == 1)]
        c.layers.append((unis, uniform_meas(n, p, rm)))
    c.draws = list(rd.random(n * depth + 1))
    return c


def haar_circuit_fast(n: int, depth: int, p: float, seed: int) -> Circuit:
    """``haar_circuit`` with batched unitary generation (its own streams:
    the unitaries are not those of ``haar_circuit`` for the same seed)."""
    ru = np.random.default_rng(seed)
    rm = np.random.default_rng(seed + 1)
    rd = np.random.default_rng(seed + 2)
    c = Circuit(n=n, depth=depth, p=p, seed=seed)
    sites = [brickwork_sites(n, t % 2 == 1) for t in range(depth)]
    mats = haar4_batch(ru, sum(len(s) for s in sites))
    pos = 0
    for t in range(depth):
        unis = [(k, mats[pos + i]) for i, k in enumerate(sites[t])]
        pos += len(sites[t])
        meas = list(np.nonzero(rm.random(n) < p)[0])
        c.layers.append((unis, [int(k) for k in meas]))
    c.draws = list(rd.random(n * depth + 1))
    return c
