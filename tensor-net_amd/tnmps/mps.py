"""Host-side mirror of tensor-net's ``MPS`` operator API over the C ABI.

Method names, argument meaning, return conventions (``None`` for an
out-of-range index, silent no-op for out-of-range unitaries) and error kinds
follow ``tensor-net/lib/mps.rs``; every method cites the reference method it
mirrors.  The state lives on the GPU; all arithmetic runs in the HIP kernels
of ``libtnmps.so``.

Truncation is given as in the reference's ``Option<BondDim<f64>>``:
``None``, ``("const", chi)`` or ``("cutoff", eps)``.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L


class MPSError(Exception):
    """Mirrors ``MPSError`` (mps.rs:126-154) plus the ABI's extra statuses."""

    KINDS = {
        L.TN_EMPTY_SYSTEM: "EmptySystem",
        L.TN_INVALID_QUANTUM_NUMBER: "InvalidQuantumNumber",
        L.TN_UNPHYSICAL_INDEX: "UnphysicalIndex",
        L.TN_STATE_INCOMPATIBLE_SHAPE: "StateIncompatibleShape",
        L.TN_OPERATOR_INCOMPATIBLE_SHAPE: "OperatorIncompatibleShape",
        L.TN_ZERO_PROB: "ZeroProbability",
        L.TN_SAMPLE_RANGE: "SampleOutOfRange",
        L.TN_CAPACITY: "Capacity",
        L.TN_DEVICE: "Device",
        L.TN_INVALID_ARG: "InvalidArgument",
        L.TN_NO_CONVERGENCE: "NoConvergence",
    }

    def __init__(self, code: int, where: str = ""):
        self.code = code
        self.kind = self.KINDS.get(code, f"status{code}")
        msg = f"{where}: {L.status_str(code)}"
        if code == L.TN_DEVICE:
            msg += f" ({L.last_error()})"
        super().__init__(msg)


def _check(code: int, where: str) -> None:
    if code != L.TN_OK:
        raise MPSError(code, where)


class MPSWarning(UserWarning):
    """A rank clamped to its bond allocation (TN_CAPACITY) or an SVD stopped
    at its sweep limit (TN_NO_CONVERGENCE) in a trajectory or batch call: the
    states are written and the call's results are returned; the handles that
    set it keep the bits (``MPS.take_flags``)."""


def _check_warn(code: int, where: str) -> None:
    """``_check`` for trajectory and batch calls: the two statuses that still
    write every state are warnings there, so a batch keeps its results."""
    if code in (L.TN_CAPACITY, L.TN_NO_CONVERGENCE):
        import warnings
        warnings.warn(MPSWarning(f"{where}: {L.status_str(code)}"), stacklevel=3)
        return
    _check(code, where)


# contraction arithmetic names (tn_set_precision)
PRECISIONS = {"f64": L.PREC_F64, "f32": L.PREC_F32, "bf16x3": L.PREC_BF16X3}


def _tune_array(rows):
    """ctypes array of tn_tune from dicts or field tuples (fields missing
    from a row take their value in TUNE_DEFAULTS)."""
    rows = list(rows)
    arr = (L.tn_tune * max(1, len(rows)))()
    for i, r in enumerate(rows):
        if isinstance(r, dict):
            vals = [r[f] if f in r else L.TUNE_DEFAULTS[f] for f in L.TUNE_FIELDS]
        else:
            vals = list(r) + [L.TUNE_DEFAULTS[f] for f in L.TUNE_FIELDS[len(r):]]
        for f, v in zip(L.TUNE_FIELDS, vals):
            setattr(arr[i], f, int(v))
    return arr, len(rows)


def to_trunc(trunc) -> L.tn_trunc:
    if trunc is None:
        return L.tn_trunc(L.TRUNC_NONE, 0, 0.0)
    kind, val = trunc
    if kind == "const":
        return L.tn_trunc(L.TRUNC_CONST, int(val), 0.0)
    if kind == "cutoff":
        return L.tn_trunc(L.TRUNC_CUTOFF, 0, float(val))
    raise ValueError(f"unknown truncation {trunc!r}")


def from_trunc(t: L.tn_trunc):
    if t.kind == L.TRUNC_NONE:
        return None
    if t.kind == L.TRUNC_CONST:
        return ("const", int(t.chi))
    return ("cutoff", float(t.eps))


def _op_buf(op, side: int) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(op, dtype=np.complex128))
    if a.shape != (side, side):
        raise MPSError(L.TN_OPERATOR_INCOMPATIBLE_SHAPE, "operator shape")
    return a.view(np.float64).reshape(-1)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _szarr(vals) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(vals, dtype=np.uintp))


def _szptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_size_t))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _meas_entries(n, msites, rit):
    """(site, forced outcome or -1, draw) per listed measurement.  An entry of
    ``msites`` is a site (``Meas::Rand``) or a (site, outcome) pair
    (``Meas::Proj``).  Draws from ``rit`` go to the ``Meas::Rand`` entries
    that ``apply_meas_prob_layer`` processes (mps.rs:2212-2219: in range,
    stably sorted by site, the first entry per site kept), in that order."""
    ops = []
    for m in msites:
        if isinstance(m, (tuple, list)):
            ops.append((int(m[0]), int(m[1])))
        else:
            ops.append((int(m), None))
    kept = {}
    for k, o in sorted((x for x in ops if x[0] < n), key=lambda x: x[0]):
        kept.setdefault(k, o)
    rmap = {k: float(next(rit)) for k in sorted(kept) if kept[k] is None}
    return [(k, -1, rmap.get(k, 0.0)) if o is None else (k, o, 0.0) for k, o in ops]


def _pack_layer(n, unis, msites, forced_layer=None, rit=None):
    """Flatten one bilayer: gate sites/ops and measurement sites/outcomes/draws.
    With ``forced_layer`` ([(k, o)]) the measurements are ``Meas::Proj``;
    otherwise as ``_meas_entries``."""
    uks = [int(k) for k, _ in unis]
    ops = [np.ascontiguousarray(np.asarray(m, dtype=np.complex128)).view(np.float64).reshape(-1)
           for _, m in unis]
    if forced_layer is not None:
        ml = [(int(k), int(o), 0.0) for (k, o) in forced_layer]
    else:
        ml = _meas_entries(n, msites, rit)
    return uks, ops, ml


class MPS:
    """Device-resident Schmidt-form MPS of qubits (``MPS<Q, C64>``, mps.rs:338-356)."""

    def __init__(self, handle: int, lib, engine: "Engine" = None):
        self._h = C.c_void_p(handle)
        self._lib = lib
        self._engine = engine  # keeps a shared engine alive while in use

    # ------------------------------------------------------------ lifecycle
    @classmethod
    def new_qubits(cls, n: int, trunc=None, chi_cap: int = 0, stream=None) -> "MPS":
        """``MPS::new_qubits`` (mps.rs:2033-2039): n qubits in |0...0>."""
        lib = L.load()
        h = C.c_void_p()
        _check(lib.tn_mps_new_qubits(int(n), to_trunc(trunc), int(chi_cap),
                                     C.c_void_p(stream) if stream else None, C.byref(h)),
               "new_qubits")
        return cls(h.value, lib)

    @classmethod
    def from_vector(cls, psi, trunc=None, chi_cap: int = 0, stream=None) -> "MPS":
        """``MPS::from_vector`` (mps.rs:516-546): the MPS of a qubit state
        vector (2^n amplitudes, qubit 0 the slowest index), normalised and
        factored site by site on the device (``factorize``, mps.rs:448-508).
        A length that is not a power of 2 is ``StateIncompatibleShape``."""
        a = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128).reshape(-1))
        n = a.size.bit_length() - 1
        if a.size == 0:
            raise MPSError(L.TN_EMPTY_SYSTEM, "from_vector")
        if a.size != 1 << n or n == 0:
            raise MPSError(L.TN_STATE_INCOMPATIBLE_SHAPE, "from_vector")
        lib = L.load()
        h = C.c_void_p()
        code = lib.tn_mps_from_vector(n, _dptr(a.view(np.float64)), to_trunc(trunc),
                                      int(chi_cap), C.c_void_p(stream) if stream else None,
                                      C.byref(h))
        if code == L.TN_CAPACITY and h.value:
            st = cls(h.value, lib)
            import warnings
            warnings.warn(MPSWarning(f"from_vector: {L.status_str(code)}"), stacklevel=2)
            return st
        _check(code, "from_vector")
        return cls(h.value, lib)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tn_mps_free(h)
            self._h = C.c_void_p()

    def free(self) -> None:
        self.__del__()

    def clone(self) -> "MPS":
        """``#[derive(Clone)]`` (mps.rs:337)."""
        h = C.c_void_p()
        _check(self._lib.tn_mps_clone(self._h, C.byref(h)), "clone")
        return MPS(h.value, self._lib)

    @property
    def n(self) -> int:
        return int(self._lib.tn_mps_n(self._h))

    def get_trunc(self):
        """``get_trunc`` (mps.rs:1216)."""
        t = L.tn_trunc()
        _check(self._lib.tn_get_trunc(self._h, C.byref(t)), "get_trunc")
        return from_trunc(t)

    def set_trunc(self, trunc) -> "MPS":
        """``set_trunc`` (mps.rs:1219-1221)."""
        _check(self._lib.tn_set_trunc(self._h, to_trunc(trunc)), "set_trunc")
        return self

    def sync(self) -> None:
        _check(self._lib.tn_sync(self._h), "sync")

    def set_precision(self, prec: str) -> "MPS":
        """Contraction arithmetic: "f64" (default), "f32" (FP32 with FP32
        accumulation) or "bf16x3" (each FP32 operand split into two BF16
        halves, three BF16 products per real product, FP32 accumulation);
        SVD, truncation and renormalisation stay FP64."""
        code = PRECISIONS[prec]
        _check(self._lib.tn_set_precision(self._h, code), "set_precision")
        return self

    def set_tuning(self, rows) -> "MPS":
        """Launch-shape table of this handle's engine (rows as for
        :meth:`Engine.set_tuning`)."""
        arr, n = _tune_array(rows)
        _check(self._lib.tn_set_tuning(self._h, arr, n), "set_tuning")
        return self

    # ------------------------------------------------------------ unitaries
    def apply_unitary1(self, k: int, op) -> "MPS":
        """``apply_unitary1`` (mps.rs:1298-1314); out-of-range k is a no-op."""
        buf = _op_buf(op, 2)
        _check(self._lib.tn_apply_u1(self._h, int(k), _dptr(buf)), "apply_unitary1")
        return self

    def apply_unitary2(self, k: int, op) -> "MPS":
        """``apply_unitary2`` (mps.rs:1332-1387); out-of-range k is a no-op."""
        buf = _op_buf(op, 4)
        _check(self._lib.tn_apply_u2(self._h, int(k), _dptr(buf)), "apply_unitary2")
        return self

    def apply_unitary2_multi(self, k: int, ops) -> "MPS":
        """``apply_unitary2_multi`` (mps.rs:1393-1411): the 4x4 operators
        multiplied in order (last one leftmost), applied at bond k as one
        ``apply_unitary2``; a no-op for k >= n - 1."""
        bufs = [_op_buf(op, 4) for op in ops]
        buf = np.ascontiguousarray(np.concatenate(bufs)) if bufs else np.zeros(32)
        _check(self._lib.tn_apply_u2_multi(self._h, int(k), len(bufs), _dptr(buf)),
               "apply_unitary2_multi")
        return self

    def apply_unitary2_noalloc(self, k: int, op) -> "MPS":
        """``apply_unitary2`` with its workspace in the library's static
        scratch (``tn_apply_u2_noalloc``): no device allocation in the call;
        bonds k-1 and k+1 at most 256."""
        buf = _op_buf(op, 4)
        _check(self._lib.tn_apply_u2_noalloc(self._h, int(k), _dptr(buf)),
               "apply_unitary2_noalloc")
        return self

    def take_flags(self) -> int:
        """The handle's sticky failure flags (``TN_FLAG_CAPACITY``,
        ``TN_FLAG_NO_CONVERGENCE``) since the last call, then cleared."""
        f = C.c_int()
        _check(self._lib.tn_take_flags(self._h, C.byref(f)), "take_flags")
        return f.value

    def apply_unis(self, unis: Sequence[Tuple[int, np.ndarray]]) -> "MPS":
        """The unitary loop of ``circuit::apply_bilayer`` (circuit.rs:637-638)
        via ``apply_uni`` (mps.rs:2093-2108)."""
        if not unis:
            return self
        ks, dims, bufs = [], [], []
        for k, mat in unis:
            a = np.asarray(mat, dtype=np.complex128)
            if a.shape not in ((2, 2), (4, 4)):
                raise MPSError(L.TN_OPERATOR_INCOMPATIBLE_SHAPE, "apply_uni")
            ks.append(int(k))
            dims.append(a.shape[0])
            bufs.append(np.ascontiguousarray(a).view(np.float64).reshape(-1))
        ksa = _szarr(ks)
        da = np.asarray(dims, dtype=np.int32)
        ops = np.ascontiguousarray(np.concatenate(bufs))
        _check(self._lib.tn_apply_unis(self._h, len(ks), _szptr(ksa), _iptr(da), _dptr(ops)),
               "apply_unis")
        return self

    # ---------------------------------------------------------- measurement
    def prob(self, k: int, p: int) -> Optional[float]:
        """``prob`` (mps.rs:1071-1074)."""
        out = C.c_double()
        rc = self._lib.tn_prob(self._h, int(k), int(p), C.byref(out))
        if rc == L.TN_NONE:
            return None
        _check(rc, "prob")
        return out.value

    def probs(self, k: int) -> Optional[List[float]]:
        """``probs`` (mps.rs:1140-1143)."""
        out = np.zeros(2)
        rc = self._lib.tn_probs(self._h, int(k), _dptr(out))
        if rc == L.TN_NONE:
            return None
        _check(rc, "probs")
        return [float(out[0]), float(out[1])]

    def measure_prob(self, k: int, r: float) -> Optional[Tuple[int, float]]:
        """``measure_prob`` (mps.rs:1479-1489); ``r`` is the uniform draw the
        reference takes from its rng."""
        o = C.c_int()
        pr = C.c_double()
        rc = self._lib.tn_measure_prob(self._h, int(k), float(r), C.byref(o), C.byref(pr))
        if rc == L.TN_NONE:
            return None
        _check(rc, "measure_prob")
        return int(o.value), pr.value

    def measure_postsel_prob(self, k: int, p: int) -> Optional[float]:
        """``measure_postsel_prob`` (mps.rs:1501-1510).  A zero-probability
        outcome raises ``MPSError`` of kind ZeroProbability (the state then
        holds NaN/inf, as the reference's does)."""
        pr = C.c_double()
        rc = self._lib.tn_measure_postsel_prob(self._h, int(k), int(p), C.byref(pr))
        if rc == L.TN_NONE:
            return None
        _check(rc, "measure_postsel_prob")
        return pr.value

    def apply_meas_prob_layer(self, meas: Sequence[Tuple[int, Optional[int], float]]):
        """``apply_meas_prob_layer`` (mps.rs:2206-2252).  ``meas`` holds
        (site, forced outcome or None for ``Meas::Rand``, uniform draw)."""
        m = len(meas)
        if m == 0:
            return []
        ks = _szarr([x[0] for x in meas])
        fo = np.asarray([-1 if x[1] is None else int(x[1]) for x in meas], dtype=np.int32)
        rs = np.asarray([float(x[2]) for x in meas], dtype=np.float64)
        cnt = C.c_size_t()
        oks = np.zeros(m, dtype=np.uintp)
        oo = np.zeros(m, dtype=np.int32)
        op = np.zeros(m, dtype=np.float64)
        _check(self._lib.tn_meas_layer(self._h, m, _szptr(ks), _iptr(fo), _dptr(rs),
                                       C.byref(cnt), _szptr(oks), _iptr(oo), _dptr(op)),
               "apply_meas_prob_layer")
        c = cnt.value
        return [(int(oks[i]), int(oo[i]), float(op[i])) for i in range(c)]

    # ---------------------------------------------------------- observables
    def entropy_vn(self, b: int) -> Optional[float]:
        """``entropy_vn`` (mps.rs:1251-1261)."""
        out = C.c_double()
        rc = self._lib.tn_entropy_vn(self._h, int(b), C.byref(out))
        if rc == L.TN_NONE:
            return None
        _check(rc, "entropy_vn")
        return out.value

    def entropy_ry_schmidt(self, a: float, b: int) -> Optional[float]:
        """``entropy_ry_schmidt`` (mps.rs:1268-1282)."""
        out = C.c_double()
        rc = self._lib.tn_entropy_ry_schmidt(self._h, float(a), int(b), C.byref(out))
        if rc == L.TN_NONE:
            return None
        _check(rc, "entropy_ry_schmidt")
        return out.value

    def bond_dims(self) -> List[int]:
        n = self.n
        if n < 2:
            return []
        out = np.zeros(n - 1, dtype=np.uintp)
        _check(self._lib.tn_bond_dims(self._h, _szptr(out)), "bond_dims")
        return [int(x) for x in out]

    def bond_dim(self, k: int) -> Optional[int]:
        """``bond_dim`` (mps.rs:1243-1245)."""
        d = self.bond_dims()
        return d[k] if 0 <= k < len(d) else None

    def max_bond_dim(self) -> Optional[int]:
        """``max_bond_dim`` (mps.rs:1227-1229)."""
        d = self.bond_dims()
        return max(d) if d else None

    def schmidt(self, b: int) -> Optional[np.ndarray]:
        """Schmidt values of bond b (``svals[b]``, mps.rs:352-353)."""
        ln = C.c_size_t()
        rc = self._lib.tn_schmidt(self._h, int(b), None, 0, C.byref(ln))
        if rc == L.TN_NONE:
            return None
        _check(rc, "schmidt")
        out = np.zeros(ln.value)
        _check(self._lib.tn_schmidt(self._h, int(b), _dptr(out), ln.value, C.byref(ln)),
               "schmidt")
        return out

    def gamma(self, k: int) -> Optional[np.ndarray]:
        """Gamma[k] (``data[k]``, axes [u_{k-1}, s_k, u_k])."""
        shp = np.zeros(3, dtype=np.uintp)
        rc = self._lib.tn_get_gamma(self._h, int(k), None, 0, _szptr(shp))
        if rc == L.TN_NONE:
            return None
        _check(rc, "gamma")
        tot = int(shp[0] * shp[1] * shp[2])
        buf = np.zeros(2 * tot)
        _check(self._lib.tn_get_gamma(self._h, int(k), _dptr(buf), buf.size, _szptr(shp)),
               "gamma")
        return buf.view(np.complex128).reshape(int(shp[0]), int(shp[1]), int(shp[2]))

    def contract(self) -> np.ndarray:
        """Flat state vector, qubit 0 slowest (``MPS::contract``, mps.rs:794-798),
        contracted on the device (``tn_contract``, n <= 30)."""
        out = np.zeros(1 << self.n, dtype=np.complex128)
        _check(self._lib.tn_contract(self._h, _dptr(out.view(np.float64)), 2 * out.size),
               "contract")
        return out

    def matrix_part(self, start: int, end: int) -> Tuple[List[Tuple[str, int]], np.ndarray]:
        """``MPS::into_matrix_part(start..end)`` (mps.rs:1926-1943): the
        reduced density matrix of sites start..end-1 (end clamped to n), the
        rest traced out through the canonical form, computed on the device
        (``tn_matrix_part``, at most 12 sites).  Returns the index list, kets
        then bras by site as the reference sorts them, and the matrix, site
        ``start`` the most significant bit of rows and columns."""
        end = min(end, self.n)
        w = end - start
        d = 1 << w if 0 < w <= 12 else 1  # other widths are refused by the call
        out = np.zeros((d, d), dtype=np.complex128)
        _check(self._lib.tn_matrix_part(self._h, start, end, _dptr(out.view(np.float64)),
                                        2 * out.size), "matrix_part")
        idx = [("Ket", k) for k in range(start, end)] + [("Bra", k) for k in range(start, end)]
        return idx, out

    def stats(self) -> Tuple[int, int]:
        u = C.c_size_t()
        s = C.c_size_t()
        _check(self._lib.tn_stats(self._h, C.byref(u), C.byref(s)), "stats")
        return u.value, s.value

    def set_timing(self, on: bool = True) -> None:
        _check(self._lib.tn_set_timing(self._h, 1 if on else 0), "set_timing")

    def perf(self) -> dict:
        p = L.tn_perf()
        _check(self._lib.tn_get_perf(self._h, C.byref(p)), "perf")
        return {name: getattr(p, name) for name, _ in L.tn_perf._fields_}

    def reset_perf(self) -> None:
        _check(self._lib.tn_reset_perf(self._h), "reset_perf")

    # ----------------------------------------------------------- circuits
    def run_circuit(self, layers, forced=None, rs=None, ent_bond: Optional[int] = None):
        """Run ``apply_bilayer`` (circuit.rs:624-659) for every layer with no
        host round trip between layers.

        ``layers``: [(unis, meas_sites)] with unis = [(k, 4x4)].  With
        ``forced`` (per layer [(k, outcome)]) the measurements replay as
        ``Meas::Proj``; otherwise they are ``Meas::Rand`` and consume the
        draws ``rs`` in processing order (sorted, deduplicated sites).
        Returns a dict like the oracle's ``run_trajectory``.
        """
        n = self.n
        depth = len(layers)
        if ent_bond is None:
            ent_bond = n // 2 - 1
        nu, uks, ops, nm, mks, fo, rr = [], [], [], [], [], [], []
        rit = iter(rs) if rs is not None else None
        for t, (unis, msites) in enumerate(layers):
            nu.append(len(unis))
            for k, mat in unis:
                uks.append(int(k))
                ops.append(np.ascontiguousarray(np.asarray(mat, dtype=np.complex128))
                           .view(np.float64).reshape(-1))
            if forced is not None:
                ml = [(int(k), int(o), 0.0) for (k, o) in forced[t]]
            else:
                srt = sorted(set(int(k) for k in msites if k < n))
                rmap = {k: next(rit) for k in srt}
                ml = [(int(k), -1, float(rmap.get(int(k), 0.0))) for k in msites]
            nm.append(len(ml))
            for (k, o, r) in ml:
                mks.append(k)
                fo.append(o)
                rr.append(r)
        nu_a = _szarr(nu)
        uks_a = _szarr(uks if uks else [0])
        ops_a = np.ascontiguousarray(np.concatenate(ops)) if ops else np.zeros(32)
        nm_a = _szarr(nm)
        mks_a = _szarr(mks if mks else [0])
        fo_a = np.asarray(fo if fo else [0], dtype=np.int32)
        rr_a = np.asarray(rr if rr else [0.0], dtype=np.float64)
        tot = max(len(mks), 1)
        ent = np.zeros(depth + 1)
        cnts = np.zeros(max(depth, 1), dtype=np.uintp)
        oo = np.zeros(tot, dtype=np.int32)
        op = np.zeros(tot, dtype=np.float64)
        _check_warn(self._lib.tn_run_circuit(self._h, depth, _szptr(nu_a), _szptr(uks_a),
                                        _dptr(ops_a), _szptr(nm_a), _szptr(mks_a),
                                        _iptr(fo_a), _dptr(rr_a), int(ent_bond), _dptr(ent),
                                        _szptr(cnts), _iptr(oo), _dptr(op)),
                     "run_circuit")
        # processed sites per layer: sorted, deduplicated, in range
        record, probs = [], []
        pos = 0
        mo = 0
        for t in range(depth):
            sites = sorted(set(k for k in mks[mo:mo + nm[t]] if k < n))
            mo += nm[t]
            c = int(cnts[t])
            assert c == len(sites)
            record.append([(sites[i], int(oo[pos + i])) for i in range(c)])
            probs.append([(sites[i], float(op[pos + i])) for i in range(c)])
            pos += c
        return {"record": record, "probs": probs, "entropy": list(ent)}


def rng_words(seed: int, c1: int, c2: int, c3: int, nblocks: int) -> np.ndarray:
    """Philox4x64-10 words of counters (b, c1, c2, c3), b < nblocks, key
    (seed, 0), from the library's host generator (``tn_rng_words``)."""
    lib = L.load()
    out = np.zeros(4 * int(nblocks), dtype=np.uint64)
    _check(lib.tn_rng_words(int(seed), int(c1), int(c2), int(c3), int(nblocks),
                            out.ctypes.data_as(C.POINTER(C.c_uint64))), "rng_words")
    return out


MASK_STREAM = 1 << 32
DRAW_STREAM = 2 << 32


class Engine:
    """A HIP stream plus the update workspace shared by many trajectories.

    Trajectories made with :meth:`new_qubits` can be advanced together with
    :meth:`batch_bilayer`: their gates and the steps of their measurement
    schedules run in shared launches.  No reference counterpart (the reference
    runs one MPS per rayon thread, workspace/src/haar_coev_probs.rs:162-203).
    """

    def __init__(self, chi_max: int, stream=None):
        self._lib = L.load()
        h = C.c_void_p()
        _check(self._lib.tn_engine_new(int(chi_max), C.c_void_p(stream) if stream else None,
                                       C.byref(h)), "engine_new")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tn_engine_free(h)
            self._h = C.c_void_p()

    def set_tuning(self, rows) -> None:
        """Replace the engine's launch-shape table: a list of dicts (or
        tuples) with the fields of ``tn_tune`` (include/tnmps.h), c_max
        increasing.  The arithmetic is the same SVD to rounding whatever the
        table says."""
        arr, n = _tune_array(rows)
        _check(self._lib.tn_engine_set_tuning(self._h, arr, n), "engine_set_tuning")

    def get_tuning(self) -> List[dict]:
        """The engine's launch-shape table as a list of dicts."""
        arr = (L.tn_tune * 16)()
        n = C.c_size_t()
        _check(self._lib.tn_engine_get_tuning(self._h, arr, 16, C.byref(n)), "engine_get_tuning")
        return [{f: getattr(arr[i], f) for f in L.TUNE_FIELDS} for i in range(n.value)]

    def stats(self) -> Tuple[int, int]:
        """(two-site updates, Jacobi sweeps) done by all handles of the engine."""
        u = C.c_size_t()
        s = C.c_size_t()
        _check(self._lib.tn_engine_stats(self._h, C.byref(u), C.byref(s)), "engine_stats")
        return u.value, s.value

    def set_timing(self, on: bool = True) -> None:
        _check(self._lib.tn_engine_set_timing(self._h, 1 if on else 0), "engine_set_timing")

    def perf(self) -> dict:
        p = L.tn_perf()
        _check(self._lib.tn_engine_get_perf(self._h, C.byref(p)), "engine_perf")
        return {name: getattr(p, name) for name, _ in L.tn_perf._fields_}

    def reset_perf(self) -> None:
        _check(self._lib.tn_engine_reset_perf(self._h), "engine_reset_perf")

    def sync(self) -> None:
        _check(self._lib.tn_engine_sync(self._h), "engine_sync")

    def new_qubits(self, n: int, trunc=None, chi_cap: int = 0) -> MPS:
        h = C.c_void_p()
        _check(self._lib.tn_mps_new_qubits_in(self._h, int(n), to_trunc(trunc), int(chi_cap),
                                              C.byref(h)), "new_qubits_in")
        return MPS(h.value, self._lib, engine=self)

    def entropies(self, states: Sequence[MPS], bonds, alphas) -> np.ndarray:
        """``entropy_ry_schmidt`` (mps.rs:1268-1282; alpha 1 is
        ``entropy_vn``) of bonds[i] of states[i], all in one launch
        (``tn_entropies``); NaN where the reference returns None."""
        cnt = len(states)
        hs = (C.c_void_p * max(cnt, 1))(*[st._h.value for st in states])
        b = _szarr(np.broadcast_to(np.asarray(bonds), (cnt,)).tolist())
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(alphas, dtype=np.float64), (cnt,)))
        out = np.zeros(max(cnt, 1))
        _check(self._lib.tn_entropies(self._h, cnt, C.cast(hs, C.POINTER(C.c_void_p)),
                                      _szptr(b), _dptr(a), _dptr(out)), "entropies")
        return out[:cnt]

    def gen_haar4(self, seed: int, ids) -> np.ndarray:
        """Haar 4x4 unitaries generated on the device (``tn_gen_haar4``):
        unitary u from counter words ``ids[u] = (c1, c2, c3)``."""
        ids = np.ascontiguousarray(np.asarray(ids, dtype=np.uint64).reshape(-1, 3))
        out = np.zeros((ids.shape[0], 4, 4), dtype=np.complex128)
        _check(self._lib.tn_gen_haar4(self._h, int(seed), ids.shape[0],
                                      ids.ctypes.data_as(C.POINTER(C.c_uint64)),
                                      _dptr(out.view(np.float64).reshape(-1))), "gen_haar4")
        return out

    def batch_bilayer_gen(self, states: Sequence[MPS], seed: int, traj_ids, layers,
                          p_meas, ent_bonds=None):
        """One generated bilayer on every state (``tn_batch_bilayer_gen``):
        layer ``layers[i]`` (or one layer for all) of trajectory
        ``traj_ids[i]``, with Haar gates made on the device and measurement
        probability ``p_meas[i]``.  Returns per state (record [(k, outcome)],
        probs [(k, p)], entropy)."""
        cnt = len(states)
        hs = (C.c_void_p * max(cnt, 1))(*[st._h.value for st in states])
        tid = np.ascontiguousarray(np.asarray(traj_ids, dtype=np.uint64).reshape(-1))
        lay = np.ascontiguousarray(np.broadcast_to(np.asarray(layers, dtype=np.uint64), (cnt,)))
        pm = np.ascontiguousarray(np.broadcast_to(np.asarray(p_meas, dtype=np.float64), (cnt,)))
        eb = _szarr(ent_bonds if ent_bonds is not None else [st.n // 2 - 1 for st in states])
        tot = max(sum(st.n for st in states), 1)
        ent = np.zeros(max(cnt, 1))
        counts = np.zeros(max(cnt, 1), dtype=np.uintp)
        sites = np.zeros(tot, dtype=np.uintp)
        oo = np.zeros(tot, dtype=np.int32)
        op = np.zeros(tot)
        _check_warn(self._lib.tn_batch_bilayer_gen(self._h, cnt, C.cast(hs, C.POINTER(C.c_void_p)),
                                              int(seed), tid.ctypes.data_as(C.POINTER(C.c_uint64)),
                                              lay.ctypes.data_as(C.POINTER(C.c_uint64)),
                                              _dptr(pm), _szptr(eb), _dptr(ent),
                                              _szptr(counts), _szptr(sites), _iptr(oo), _dptr(op)),
               "batch_bilayer_gen")
        out, pos = [], 0
        for i in range(cnt):
            c = int(counts[i])
            rec = [(int(sites[pos + j]), int(oo[pos + j])) for j in range(c)]
            prs = [(int(sites[pos + j]), float(op[pos + j])) for j in range(c)]
            pos += c
            out.append((rec, prs, float(ent[i])))
        return out

    def batch_bilayer(self, states: Sequence[MPS], layers, forced=None, draws=None,
                      ent_bonds=None):
        """One bilayer on every state, in lock step (``tn_batch_bilayer``).

        ``layers[i] = (unis, meas_sites)`` for state i; ``forced[i]`` is its
        [(k, outcome)] record (``Meas::Proj``) or ``draws[i]`` an iterator of
        uniform draws (``Meas::Rand``).  Returns per state a tuple
        (record [(k, outcome)], probs [(k, p)], entropy at ent_bonds[i]).
        """
        cnt = len(states)
        nu, uks, ops, nm, mks, fo, rr, sites = [], [], [], [], [], [], [], []
        for i, st in enumerate(states):
            unis, msites = layers[i]
            u, o, ml = _pack_layer(st.n, unis, msites,
                                   forced[i] if forced is not None else None,
                                   draws[i] if draws is not None else None)
            nu.append(len(u))
            uks += u
            ops += o
            nm.append(len(ml))
            for (k, f, r) in ml:
                mks.append(k)
                fo.append(f)
                rr.append(r)
            sites.append(sorted(set(k for k, _, _ in ml if k < st.n)))
        hs = (C.c_void_p * cnt)(*[st._h.value for st in states])
        nu_a, nm_a = _szarr(nu), _szarr(nm)
        uks_a = _szarr(uks if uks else [0])
        ops_a = np.ascontiguousarray(np.concatenate(ops)) if ops else np.zeros(32)
        mks_a = _szarr(mks if mks else [0])
        fo_a = np.asarray(fo if fo else [0], dtype=np.int32)
        rr_a = np.asarray(rr if rr else [0.0], dtype=np.float64)
        eb = _szarr(ent_bonds if ent_bonds is not None else [st.n // 2 - 1 for st in states])
        ent = np.zeros(cnt)
        counts = np.zeros(cnt, dtype=np.uintp)
        tot = max(len(mks), 1)
        oo = np.zeros(tot, dtype=np.int32)
        op = np.zeros(tot)
        _check_warn(self._lib.tn_batch_bilayer(self._h, cnt, C.cast(hs, C.POINTER(C.c_void_p)),
                                          _szptr(nu_a), _szptr(uks_a), _dptr(ops_a),
                                          _szptr(nm_a), _szptr(mks_a), _iptr(fo_a),
                                          _dptr(rr_a), _szptr(eb), _dptr(ent), _szptr(counts),
                                          _iptr(oo), _dptr(op)),
               "batch_bilayer")
        out, pos = [], 0
        for i in range(cnt):
            c = int(counts[i])
            assert c == len(sites[i])
            rec = [(sites[i][j], int(oo[pos + j])) for j in range(c)]
            prs = [(sites[i][j], float(op[pos + j])) for j in range(c)]
            pos += c
            out.append((rec, prs, float(ent[i])))
        return out
