"""Literal CPU restatement of tensor-net's Schmidt-form MPS (test oracle only).

Every method follows the cited lines of ``tensor-net/lib/mps.rs`` (reference
paths are relative to the reference checkout).  Arithmetic is complex128 /
float64 numpy; the SVD is LAPACK ``zgesvd`` (scipy's OpenBLAS build) with
full U and V^H, which is what ``ndarray-linalg``'s ``svd_into(true, true)``
calls (mps.rs:299).

Storage conventions are the reference's:
  * ``G[k]``  complex128 array of shape (chiL, 2, chiR), axes [u_{k-1}, s_k, u_k]
    (mps.rs:345-350).
  * ``L[k]``  float64 Schmidt values on bond k (between sites k and k+1),
    k = 0 .. n-2 (mps.rs:352-353).
  * truncation ``trunc``: None | ("const", chi) | ("cutoff", eps)
    (``BondDim``, mps.rs:157-166).

The oracle is test infrastructure; see the package docstring.
"""
from __future__ import annotations

import math
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np
import scipy.linalg as sla

F64_EPS = float(np.finfo(np.float64).eps)        # f64::EPSILON
F64_MIN_POSITIVE = float(np.finfo(np.float64).tiny)  # smallest normal f64


class OracleError(Exception):
    """Mirrors ``MPSError`` (mps.rs:126-154); ``kind`` is the variant name."""

    def __init__(self, kind: str):
        super().__init__(kind)
        self.kind = kind


def _is_normal(x: float) -> bool:
    """``f64::is_normal``: finite, non-zero, not subnormal."""
    return math.isfinite(x) and x != 0.0 and abs(x) >= F64_MIN_POSITIVE


def leading_count(s: np.ndarray, thresh: float) -> int:
    """``take_while(|sj| sj.is_normal() && sj > thresh).count()`` (mps.rs:254-257)."""
    n = 0
    for v in s:
        if _is_normal(float(v)) and v > thresh:
            n += 1
        else:
            break
    return n


def truncation_rank(s_normalised: np.ndarray, trunc) -> int:
    """Rank rule of ``process_svd`` (mps.rs:250-274).

    ``Const(r)``: min(max(r, 1), #leading normal values > f64::EPSILON)
    ``Cutoff(e)``: max(1, #leading normal values > |e|)
    ``None``: same as Cutoff(f64::EPSILON).
    """
    if trunc is None:
        return max(leading_count(s_normalised, F64_EPS), 1)
    kind, val = trunc
    if kind == "const":
        n = leading_count(s_normalised, F64_EPS)
        return min(max(int(val), 1), n)
    if kind == "cutoff":
        return max(leading_count(s_normalised, abs(float(val))), 1)
    raise ValueError(f"unknown truncation kind {kind!r}")


def process_svd(u: np.ndarray, s: np.ndarray, q: np.ndarray, trunc):
    """mps.rs:236-293: normalise, pick rank, renormalise, scale rows of q.

    Returns (u[:, :r], s[:r] renormalised, q[:r, :] * (s_j * norm2), r).
    """
    s = np.array(s, dtype=np.float64)
    norm = math.sqrt(sum(float(x) * float(x) for x in s))       # :243-247
    s = s / norm                                                  # :248
    rank = truncation_rank(s, trunc)                              # :249-274
    s = s[:rank].copy()                                           # :275-276
    norm2 = math.sqrt(sum(float(x) * float(x) for x in s))       # :277-281
    s = s / norm2                                                 # :282
    q = q[:rank, :].copy()                                        # :283
    q *= (s * norm2)[:, None]                                     # :284-290
    u = u[:, :rank].copy()                                        # :291
    return u, s, q, rank


def local_decomp(theta: np.ndarray, trunc, driver: str = "gesvd"):
    """``SchmidtDecomp::local_decomp`` (mps.rs:295-303): zgesvd(A, A) + process_svd.

    ``driver="gesdd"`` takes LAPACK's divide-and-conquer zgesdd instead: the
    same SVD with different rounding, about 15 times faster on 512 x 512
    thetas.  Only the full-size tests use it (their tolerances do not depend
    on the reference's rounding); tests/test_oracle_pins.py checks it against
    the zgesvd oracle."""
    u, s, vh = sla.svd(theta, full_matrices=True, lapack_driver=driver,
                       check_finite=False)
    return process_svd(u, s, vh, trunc)


class VidalMPS:
    """Restatement of ``MPS<Q, C64>`` for qubits (mps.rs:338-356)."""

    svd_driver = "gesvd"  # LAPACK driver of local_decomp (see there)

    def __init__(self, n: int, trunc=None):
        """``MPS::new_qubits`` (mps.rs:2033-2039) -> ``MPS::new`` (mps.rs:370-391)."""
        if n < 1:
            raise OracleError("EmptySystem")                      # :375
        self.n = n
        self.G: List[np.ndarray] = []
        for _ in range(n):
            g = np.zeros((1, 2, 1), dtype=np.complex128)
            g[0, 0, 0] = 1.0                                      # :382
            self.G.append(g)
        self.L: List[np.ndarray] = [np.array([1.0]) for _ in range(n - 1)]  # :385-388
        self.trunc = trunc

    # ------------------------------------------------------------------ copy
    def copy(self) -> "VidalMPS":
        m = VidalMPS.__new__(VidalMPS)
        m.n = self.n
        m.G = [g.copy() for g in self.G]
        m.L = [l.copy() for l in self.L]
        m.trunc = self.trunc
        m.svd_driver = self.svd_driver
        return m

    # ----------------------------------------------------------- contraction
    @staticmethod
    def do_contract2(l: np.ndarray, s: np.ndarray, r: np.ndarray) -> np.ndarray:
        """mps.rs:628-651: (chiL*2, chiM) . diag(s) (chiM, 2*chiR)."""
        shl = l.shape
        lm = l.reshape(shl[0] * shl[1], shl[2])
        shr = r.shape
        rm = r.reshape(shr[0], shr[1] * shr[2]) * s[:, None]      # :644-649
        return lm @ rm                                            # :650 (zgemm)

    def contract_local2(self, k: int) -> np.ndarray:
        """mps.rs:732-754: theta with lambda[k-1] on rows, lambda[k+1] on columns."""
        shk = self.G[k].shape
        shkp1 = self.G[k + 1].shape
        q = self.do_contract2(self.G[k], self.L[k], self.G[k + 1])
        q = q.reshape(shk[0], shk[1] * shkp1[1], shkp1[2])        # do_contract3 :660-673
        if k != 0:
            q = q * self.L[k - 1][:, None, None]                  # :737-744
        if k != self.n - 2:
            q = q * self.L[k + 1][None, None, :]                  # :745-752
        return q.reshape(shk[0] * shk[1], shkp1[1] * shkp1[2])

    def contract_local3(self, k: int) -> np.ndarray:
        """mps.rs:767-773."""
        shk = self.G[k].shape
        shkp1 = self.G[k + 1].shape
        return self.contract_local2(k).reshape(shk[0], shk[1] * shkp1[1], shkp1[2])

    def contract(self) -> np.ndarray:
        """Flat state vector (mps.rs:794-798), qubit 0 the slowest index.

        Sequential left-to-right contraction; the reference's binary split
        (mps.rs:682-700) is the same product in a different association order.
        """
        acc = self.G[0].reshape(2, -1)                            # (1*2, chi0)
        for k in range(1, self.n):
            g = self.G[k] * self.L[k - 1][:, None, None]
            acc = acc @ g.reshape(g.shape[0], -1)
            acc = acc.reshape(-1, g.shape[2])
        return acc.reshape(-1)

    # ------------------------------------------------------- split + renorm
    def _split(self, k: int, u: np.ndarray, s: np.ndarray, q: np.ndarray, rank: int,
               shk, shkp1) -> None:
        """Shared tail of local_refactor (mps.rs:852-880) / apply_unitary2 (:1357-1385)."""
        gk_new = u.reshape(shk[0], shk[1], rank)
        if k != 0:
            gk_new = gk_new / self.L[k - 1][:, None, None]        # :855-862
        gkp1_new = q.reshape(rank, shkp1[1], shkp1[2])
        gkp1_new = gkp1_new / s[:, None, None]                    # :864-869
        if k != self.n - 2:
            gkp1_new = gkp1_new / self.L[k + 1][None, None, :]    # :870-876
        self.G[k] = np.ascontiguousarray(gk_new)
        self.L[k] = s
        self.G[k + 1] = np.ascontiguousarray(gkp1_new)
        self.local_renormalize(k)                                 # :879
        self.local_renormalize(k + 1)                             # :880

    def local_refactor(self, k: int) -> None:
        """mps.rs:846-880: contract -> SVD -> split -> renormalise, no gate."""
        shk = self.G[k].shape
        shkp1 = self.G[k + 1].shape
        u, s, q, rank = local_decomp(self.contract_local2(k), self.trunc, self.svd_driver)
        self._split(k, u, s, q, rank, shk, shkp1)

    def refactor_sweep(self) -> None:
        """mps.rs:894-897: L->R over 0..n-1 then R->L over (0..n-2).rev()."""
        for k in range(0, self.n - 1):
            self.local_refactor(k)
        for k in reversed(range(0, self.n - 2)):
            self.local_refactor(k)

    # ------------------------------------------------------------ unitaries
    def apply_unitary1(self, k: int, op: np.ndarray) -> "VidalMPS":
        """mps.rs:1298-1314.  Out-of-range k is a silent no-op."""
        if k >= self.n:
            return self                                           # :1301
        op = np.asarray(op, dtype=np.complex128)
        if op.shape != (2, 2):
            raise OracleError("OperatorIncompatibleShape")        # :1303
        if k == self.n - 1:
            assert self.G[k].shape[2] == 1                        # :1305
            gmat = self.G[k][:, :, 0]
            self.G[k][:, :, 0] = gmat @ op.T                      # :1306-1308
        else:
            self.G[k] = np.einsum("ts,vsu->vtu", op, self.G[k])   # :1310-1311
        return self

    def apply_unitary2(self, k: int, op: np.ndarray) -> "VidalMPS":
        """mps.rs:1332-1387.  Qubit k is the most significant bit of the 4x4."""
        if self.n == 1 or k >= self.n - 1:
            return self                                           # :1335
        shk = self.G[k].shape
        shkp1 = self.G[k + 1].shape
        op = np.asarray(op, dtype=np.complex128)
        if op.shape != (shk[1] * shkp1[1], shk[1] * shkp1[1]):
            raise OracleError("OperatorIncompatibleShape")        # :1339-1341
        q = self.contract_local3(k)
        if k == self.n - 2:
            assert q.shape[2] == 1                                # :1345
            qmat = q[:, :, 0]
            q[:, :, 0] = qmat @ op.T                              # :1346-1349
        else:
            q = np.einsum("ts,vsw->vtw", op, q)                   # :1350-1351
        u, s, qq, rank = local_decomp(
            q.reshape(shk[0] * shk[1], shkp1[1] * shkp1[2]), self.trunc,
            self.svd_driver)                                      # :1354-1356
        self._split(k, u, s, qq, rank, shk, shkp1)               # :1357-1385
        return self

    # ------------------------------------------------------ probabilities
    def local_prob(self, k: int, p: int) -> float:
        """mps.rs:1027-1066."""
        if k == 0:
            g = self.G[k][0, p, :]
            return float(np.sum((g.conj() * g).real * self.L[k] ** 2))
        if k == self.n - 1:
            g = self.G[k][:, p, 0]
            return float(np.sum((g.conj() * g).real * self.L[k - 1] ** 2))
        g = self.G[k][:, p, :]
        inner = np.sum((g.conj() * g).real * (self.L[k] ** 2)[None, :], axis=1)
        return float(np.sum(inner * self.L[k - 1] ** 2))

    def local_probs(self, k: int) -> List[float]:
        """mps.rs:1079-1134."""
        if self.n == 1:
            return [float((g.conj() * g).real) for g in self.G[k].reshape(-1)]
        return [self.local_prob(k, s) for s in range(2)]

    def prob(self, k: int, p: int) -> Optional[float]:
        """mps.rs:1071-1074: None when k or p is out of range."""
        if k >= self.n or p >= 2:
            return None
        return self.local_prob(k, p)

    def probs(self, k: int) -> Optional[List[float]]:
        """mps.rs:1140-1143."""
        if k >= self.n:
            return None
        return self.local_probs(k)

    def local_norm(self, k: int) -> complex:
        """mps.rs:1147-1202 (complex sqrt of a real-valued complex sum)."""
        g = self.G[k]
        if k == 0:
            g0 = g[0, :, :]
            acc = np.sum(g0.conj() * g0 * (self.L[k] ** 2)[None, :])
        elif k == self.n - 1:
            g0 = g[:, :, 0]
            acc = np.sum(np.sum(g0.conj() * g0, axis=1) * self.L[k - 1] ** 2)
        else:
            inner = np.sum(g.conj() * g * (self.L[k] ** 2)[None, None, :], axis=(1, 2))
            acc = np.sum(inner * self.L[k - 1] ** 2)
        return complex(np.sqrt(complex(acc)))

    def local_renormalize(self, k: int) -> None:
        """mps.rs:1206-1209."""
        self.G[k] = self.G[k] / self.local_norm(k)

    # ---------------------------------------------------------- measurement
    def sample_state(self, k: int, r: float) -> Tuple[int, float]:
        """mps.rs:1427-1447 with the uniform draw ``r`` supplied by the caller.

        The reference unwraps ``find_map``; an ``r`` at or above the total
        probability panics there.  The oracle raises instead.
        """
        probs = self.local_probs(k)
        cu = 0.0
        for j, pr in enumerate(probs):
            cu += pr
            if r < cu:
                return j, probs[j]
        raise OracleError("SampleOutOfRange")

    def project(self, k: int, p: int, renorm: complex) -> None:
        """mps.rs:1453-1465: keep slice p divided by renorm, zero the others."""
        g = self.G[k]
        for j in range(g.shape[1]):
            if j == p:
                g[:, j, :] = g[:, j, :] / renorm
            else:
                g[:, j, :] = 0.0

    def measure_prob(self, k: int, r: float) -> Optional[Tuple[int, float]]:
        """mps.rs:1479-1489 with the uniform draw supplied."""
        if k >= self.n:
            return None
        p, prob = self.sample_state(k, r)
        self.project(k, p, complex(math.sqrt(prob)))
        self.refactor_sweep()
        return p, prob

    def measure_postsel_prob(self, k: int, p: int) -> Optional[float]:
        """mps.rs:1501-1510: forced outcome; zero probability gives NaN/inf as there."""
        if k >= self.n or p >= 2:
            return None
        prob = self.local_prob(k, p)
        with np.errstate(divide="ignore", invalid="ignore"):
            self.project(k, p, complex(math.sqrt(prob)))
        self.refactor_sweep()
        return prob

    def apply_meas_prob_layer(self, meas: Sequence[Tuple[int, Optional[int], float]]):
        """mps.rs:2206-2252.

        ``meas`` holds (site, forced_outcome_or_None, r): ``None`` is
        ``Meas::Rand`` (consumes ``r``), an int is ``Meas::Proj``.  Sites out
        of range are dropped, the rest stably sorted by site and deduplicated
        keeping the first (``sort_by_key`` + ``dedup_by_key``).
        Returns [(site, outcome, prob)] in processing order.
        """
        ms = [m for m in meas if m[0] < self.n]                   # :2212-2216
        if not ms:
            return []                                             # :2217
        ms.sort(key=lambda m: m[0])                               # :2218
        dedup = []
        for m in ms:                                              # :2219
            if dedup and dedup[-1][0] == m[0]:
                continue
            dedup.append(m)
        last = None
        res = []
        for (k, forced, r) in dedup:
            if last is not None:
                for j in range(last, k):                          # :2225-2228
                    self.local_refactor(j)
            if forced is None:
                p, prob = self.sample_state(k, r)                 # :2230-2233
            else:
                p = int(forced)
                prob = self.local_prob(k, p)                      # :2236-2240
            with np.errstate(divide="ignore", invalid="ignore"):
                self.project(k, p, complex(math.sqrt(prob)))
            res.append((k, p, prob))
            last = k
        if last is not None:
            for j in range(last, self.n - 1):                     # :2246-2249
                self.local_refactor(j)
        for j in reversed(range(0, self.n - 2)):                  # :2250
            self.local_refactor(j)
        return res

    # ---------------------------------------------------------- observables
    def entropy_vn(self, b: int) -> Optional[float]:
        """mps.rs:1251-1261: -sum s^2 ln s^2 over lambda[b]."""
        if b >= len(self.L):
            return None
        acc = 0.0
        for sk in self.L[b]:
            sk2 = float(sk) * float(sk)
            acc += -sk2 * math.log(sk2)
        return acc

    def entropy_ry_schmidt(self, a: float, b: int) -> Optional[float]:
        """mps.rs:1268-1282."""
        if a == 1.0:
            return self.entropy_vn(b)
        if b >= len(self.L):
            return None
        acc = sum((float(sk) * float(sk)) ** a for sk in self.L[b])
        return math.log(acc) / (1.0 - a)

    def matrix_part(self, a: int, b: int) -> np.ndarray:
        """``MPS::into_matrix_part(a..b)`` (mps.rs:1926-1943) of the network
        ``into_network_part`` builds (mps.rs:1742-1829): Gamma of sites a..b-1
        with lambda[k] on the bonds inside the range (:1778-1791), the network
        and its conjugate, and diag(lambda^2) tracers on bond a-1 (:1801-1813)
        and on bond b-1 (:1814-1826), which trace out the sites outside the
        range through the canonical form.  ``part.end`` is clamped to n
        (:1746).  Indices sorted kets (rows) before bras, each by site, so
        site a is the most significant bit of the row index (:1930-1942)."""
        b = min(b, self.n)
        if not 0 <= a < b:
            raise OracleError("InvalidArgument")
        chil = self.G[a].shape[0]
        t = np.eye(chil, dtype=np.complex128)                     # (alpha, alpha')
        for k in range(a, b):
            g = self.G[k]
            if k > a:
                g = g * self.L[k - 1][:, None, None]              # lambda inside the range
            t = (t @ g.reshape(g.shape[0], -1)).reshape(-1, g.shape[2])
        d = 2 ** (b - a)
        t = t.reshape(chil, d, -1)                                # (alpha, i, beta)
        wl = self.L[a - 1] ** 2 if a > 0 else np.ones(1)
        wr = self.L[b - 1] ** 2 if b < self.n else np.ones(1)
        return np.einsum("a,b,aib,ajb->ij", wl, wr, t, t.conj())

    def bond_dim(self, k: int) -> Optional[int]:
        """mps.rs:1243-1245."""
        return len(self.L[k]) if k < len(self.L) else None

    def bond_dims(self) -> List[int]:
        return [len(l) for l in self.L]

    def max_bond_dim(self) -> Optional[int]:
        """mps.rs:1227-1229."""
        return max(self.bond_dims()) if self.L else None


def from_vector(n: int, psi, trunc=None, driver: str = "gesvd") -> VidalMPS:
    """``MPS::from_vector`` (mps.rs:516-546) -> ``factorize`` (mps.rs:448-508)
    for n qubits: normalise (:531-537), then for each site fuse its index with
    the previous Schmidt index (:465-470), decompose (local_decomp, :474),
    Gamma[k] = U[:, :r] with rows divided by the previous Schmidt values
    (:478-492), q <- the scaled right factor; the last Gamma is the final q
    with the same division (:499-505).  The SVD is the economy-size one: the
    rows of V^H beyond min(m, n) that zgesvd's full form adds are never kept
    (process_svd keeps r <= min(m, n) of them)."""
    psi = np.asarray(psi, dtype=np.complex128).reshape(-1)
    if n < 1:
        raise OracleError("EmptySystem")
    if psi.size != 2 ** n:
        raise OracleError("StateIncompatibleShape")
    norm = math.sqrt(sum(float(abs(a)) ** 2 for a in psi))          # :531-536
    psi = psi / norm
    st = VidalMPS(n, trunc)
    if n == 1:
        st.G = [psi.reshape(1, 2, 1).copy()]                        # :538-544
        return st
    G: List[np.ndarray] = []
    L: List[np.ndarray] = []
    udim = 1
    q = psi.reshape(1, -1)
    for _ in range(n - 1):
        m = udim * 2                                                 # :465-470
        q = q.reshape(m, q.size // m)
        u, s, vh = sla.svd(q, full_matrices=False, lapack_driver=driver,
                           check_finite=False)
        u, s, q, rank = process_svd(u, s, vh, trunc)                 # :474
        g = u[:, :rank].reshape(udim, 2, rank)                       # :478-483
        if L:
            g = g / L[-1][:, None, None]                             # :484-490
        G.append(np.ascontiguousarray(g))
        L.append(s)
        udim = rank
    G.append(np.ascontiguousarray(q.reshape(udim, 2, 1) / L[-1][:, None, None]))  # :499-505
    st.G, st.L = G, L
    return st


def apply_bilayer(state: VidalMPS, unis, meas):
    """``circuit::apply_bilayer`` (circuit.rs:624-659).

    ``unis``: [(k, 4x4 or 2x2 matrix)] applied in order (``apply_uni``,
    mps.rs:2093-2108); ``meas``: as ``VidalMPS.apply_meas_prob_layer``.
    Returns [(site, outcome, prob)].
    """
    for k, mat in unis:
        mat = np.asarray(mat)
        if mat.shape == (2, 2):
            state.apply_unitary1(k, mat)
        elif mat.shape == (4, 4):
            state.apply_unitary2(k, mat)
        else:
            raise OracleError("OperatorIncompatibleShape")        # mps.rs:2104
    return state.apply_meas_prob_layer(meas)


def run_trajectory(n: int, layers, trunc, rs: Optional[Iterable[float]] = None,
                   forced: Optional[Sequence[Sequence[Tuple[int, int]]]] = None,
                   entropy_bond: Optional[int] = None):
    """Drive one trajectory through ``apply_bilayer`` per layer.

    ``layers``: [(unis, meas_sites)].  With ``forced`` given, each layer's
    measurements replay that record as ``Meas::Proj`` (the reference's own
    replay protocol, haar_coev_probs.rs:60-79); otherwise ``Meas::Rand`` draws
    consume ``rs`` in processing order.  Returns dict with the outcome record,
    probabilities, the entropy at ``entropy_bond`` before layer 1 and after
    each layer (circuit_old.rs:452-454, 502-503), and the final state.
    """
    st = VidalMPS(n, trunc)
    if entropy_bond is None:
        entropy_bond = n // 2 - 1
    rit = iter(rs) if rs is not None else None
    record, probs, ent = [], [], [st.entropy_vn(entropy_bond)]
    for t, (unis, msites) in enumerate(layers):
        if forced is not None:
            mlist = [(k, o, 0.0) for (k, o) in forced[t]]
        else:
            mlist = [(k, None, 0.0) for k in msites]
            # draws are consumed in processing (sorted, deduplicated) order
            srt = sorted(set(k for k in msites if k < n))
            rmap = {k: next(rit) for k in srt}
            mlist = [(k, None, rmap.get(k, 0.0)) for k in msites]
        res = apply_bilayer(st, unis, mlist)
        record.append([(k, o) for (k, o, _) in res])
        probs.append([(k, pr) for (k, _, pr) in res])
        ent.append(st.entropy_vn(entropy_bond))
    return {"record": record, "probs": probs, "entropy": ent, "state": st}
