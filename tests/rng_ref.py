"""CPU restatement of the library's counter-based workload generator
(declared in include/tnmps.h, kernels in tensor-net_amd/csrc/tn_rng.h and
k_haar4), for the tests only: Philox4x64-10 on Python integers, Box-Muller
normals, and a Haar unitary as numpy's QR with R's diagonal phases divided
out of Q's columns (the construction of tensor-net/lib/gate.rs:888-909)."""
import math

import numpy as np

M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
MASK64 = (1 << 64) - 1
MASK_STREAM = 1 << 32
DRAW_STREAM = 2 << 32


def philox(c, k):
    """Philox4x64-10 (Salmon et al., SC 2011) of counter c (4 words), key k (2)."""
    c0, c1, c2, c3 = (int(x) & MASK64 for x in c)
    k0, k1 = (int(x) & MASK64 for x in k)
    for _ in range(10):
        p0, p1 = M0 * c0, M1 * c2
        c0, c1, c2, c3 = (p1 >> 64) ^ c1 ^ k0, p1 & MASK64, (p0 >> 64) ^ c3 ^ k1, p0 & MASK64
        k0, k1 = (k0 + W0) & MASK64, (k1 + W1) & MASK64
    return [c0, c1, c2, c3]


def words(seed, c1, c2, c3, nblocks):
    out = []
    for b in range(nblocks):
        out += philox((b, c1, c2, c3), (seed, 0))
    return out


def u01(w):
    return (int(w) >> 11) * 2.0 ** -53


def haar4(seed, c1, c2, c3):
    w = words(seed, c1, c2, c3, 8)
    z = np.zeros((4, 4), dtype=np.complex128)
    for b in range(8):
        for h in range(2):
            rad = math.sqrt(-2.0 * math.log(1.0 - u01(w[4 * b + 2 * h])))
            ang = 2.0 * math.pi * u01(w[4 * b + 2 * h + 1])
            e = 2 * b + h
            z[e // 4, e % 4] = complex(rad * math.cos(ang), rad * math.sin(ang))
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * np.conj(d / np.abs(d))[None, :]


def mask(seed, traj, layer, n, p):
    w = words(seed, MASK_STREAM, layer, traj, (n + 3) // 4)
    return [q for q in range(n) if u01(w[q]) < p]


def draws(seed, traj, layer, m):
    w = words(seed, DRAW_STREAM, layer, traj, (m + 3) // 4)
    return [u01(x) for x in w[:m]]
