"""Circuit, measurement-record and manifest files of the reference's drivers,
read and written in the reference's layouts, so circuit sets and outcome
records made by either side replay on the other.

* serde_cbor 0.11 files (``save_cbor`` / ``load_cbor``,
  tensor-net/lib/circuit.rs:476-517): a unitary circuit is ``Vec<Vec<Uni>>``
  (circuit.rs:209-215), a measurement layer list ``Vec<Vec<Meas>>``
  (circuit.rs:344-348), and a circuit set is described by ``MiptManifest``
  (workspace/lib/haar.rs:211-320).
* npz files: unitary layers as ``depth``/``targs``/``mats``
  (``save_unis``/``load_unis``, workspace/lib/haar.rs:54-100) and measurement
  layers per p as ``pvals``/``layer_lens``/``targs``/``ops``
  (``save_meas``/``load_meas``, workspace/lib/lib.rs:61-147).

The CBOR codec follows RFC 8949.  Values map to serde's data model the way
serde_cbor 0.11.2 does by default: structs are maps keyed by field name,
enums are externally tagged (``{"Variant": payload}``, a unit variant is its
name as a text string), tuples and sequences are arrays, and a float is
written in the shortest of half, single and double precision that holds it
exactly.  ``ndarray`` arrays (0.15, feature ``serde``) are the struct
``{"v": 1, "dim": [rows, cols], "data": [...]}`` with row-major data, and a
``Complex64`` (num-complex 0.4) is the pair ``[re, im]``.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass
from typing import Any, Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np

__all__ = [
    "CborError", "cbor_dumps", "cbor_loads",
    "Mat", "Gate", "Rand", "Proj", "Uni", "Meas",
    "gate_matrix", "uni_to_serde", "uni_from_serde", "meas_to_serde", "meas_from_serde",
    "save_cbor", "load_cbor", "save_uni_circuit", "load_uni_circuit",
    "save_meas_layers", "load_meas_layers", "MiptManifest",
    "save_unis_npz", "load_unis_npz", "save_meas_npz", "load_meas_npz",
]


# ---------------------------------------------------------------- CBOR codec
class CborError(ValueError):
    """Malformed or unsupported CBOR input."""


_MAX_DEPTH = 512


class _Reader:
    def __init__(self, data: bytes):
        self.b = memoryview(bytes(data))
        self.i = 0

    def take(self, n: int) -> bytes:
        if n < 0 or self.i + n > len(self.b):
            raise CborError("truncated CBOR input")
        out = bytes(self.b[self.i:self.i + n])
        self.i += n
        return out

    def arg(self, ai: int) -> Optional[int]:
        """The argument of a head with additional information ai (None for an
        indefinite length)."""
        if ai < 24:
            return ai
        if ai == 24:
            return self.take(1)[0]
        if ai == 25:
            return struct.unpack(">H", self.take(2))[0]
        if ai == 26:
            return struct.unpack(">I", self.take(4))[0]
        if ai == 27:
            return struct.unpack(">Q", self.take(8))[0]
        if ai == 31:
            return None
        raise CborError(f"reserved additional information {ai}")


_BREAK = object()


def _item(rd: _Reader, depth: int) -> Any:
    if depth > _MAX_DEPTH:
        raise CborError("CBOR nesting too deep")
    ib = rd.take(1)[0]
    major, ai = ib >> 5, ib & 31
    if major == 7:
        if ai == 20:
            return False
        if ai == 21:
            return True
        if ai in (22, 23):
            return None
        if ai == 25:
            return float(np.frombuffer(rd.take(2), dtype=">f2")[0])
        if ai == 26:
            return struct.unpack(">f", rd.take(4))[0]
        if ai == 27:
            return struct.unpack(">d", rd.take(8))[0]
        if ai == 31:
            return _BREAK
        if ai == 24:
            rd.take(1)
            return None
        if ai < 20:
            return None
        raise CborError(f"reserved simple value {ai}")
    n = rd.arg(ai)
    if major == 0:
        if n is None:
            raise CborError("indefinite-length integer")
        return n
    if major == 1:
        if n is None:
            raise CborError("indefinite-length integer")
        return -1 - n
    if major in (2, 3):
        if n is None:
            parts = []
            while True:
                x = _item(rd, depth + 1)
                if x is _BREAK:
                    break
                want = bytes if major == 2 else str
                if not isinstance(x, want):
                    raise CborError("bad chunk in an indefinite-length string")
                parts.append(x)
            return b"".join(parts) if major == 2 else "".join(parts)
        raw = rd.take(n)
        if major == 2:
            return raw
        try:
            return raw.decode("utf-8")
        except UnicodeDecodeError as e:
            raise CborError("invalid UTF-8 in a text string") from e
    if major == 4:
        out = []
        if n is None:
            while True:
                x = _item(rd, depth + 1)
                if x is _BREAK:
                    return out
                out.append(x)
        for _ in range(n):
            x = _item(rd, depth + 1)
            if x is _BREAK:
                raise CborError("unexpected break")
            out.append(x)
        return out
    if major == 5:
        out = {}
        count = 0
        while n is None or count < n:
            key = _item(rd, depth + 1)
            if key is _BREAK:
                if n is None:
                    return out
                raise CborError("unexpected break")
            val = _item(rd, depth + 1)
            if val is _BREAK:
                raise CborError("unexpected break")
            out[_hashable(key)] = val
            count += 1
        return out
    # major 6: tag; bignums (2, 3) become ints, other tags are transparent
    if n is None:
        raise CborError("indefinite-length tag")
    x = _item(rd, depth + 1)
    if n in (2, 3) and isinstance(x, bytes):
        v = int.from_bytes(x, "big")
        return v if n == 2 else -1 - v
    return x


def _hashable(x: Any) -> Any:
    if isinstance(x, list):
        return tuple(_hashable(v) for v in x)
    if isinstance(x, dict):
        return tuple((k, _hashable(v)) for k, v in x.items())
    return x


def cbor_loads(data: bytes) -> Any:
    """Decode one CBOR data item that spans all of ``data``."""
    rd = _Reader(data)
    x = _item(rd, 0)
    if x is _BREAK:
        raise CborError("unexpected break")
    if rd.i != len(rd.b):
        raise CborError(f"{len(rd.b) - rd.i} trailing bytes after the CBOR item")
    return x


def _head(major: int, n: int) -> bytes:
    if n < 24:
        return bytes([(major << 5) | n])
    if n < 1 << 8:
        return bytes([(major << 5) | 24, n])
    if n < 1 << 16:
        return bytes([(major << 5) | 25]) + struct.pack(">H", n)
    if n < 1 << 32:
        return bytes([(major << 5) | 26]) + struct.pack(">I", n)
    if n < 1 << 64:
        return bytes([(major << 5) | 27]) + struct.pack(">Q", n)
    raise CborError("integer out of the 64-bit range")


def _float(v: float) -> bytes:
    """serde_cbor's float rule: double -> single when exact, single -> half
    when exact; infinities and NaN as half."""
    if math.isnan(v):
        return b"\xf9\x7e\x00"
    if math.isinf(v):
        return b"\xf9\x7c\x00" if v > 0 else b"\xf9\xfc\x00"
    with np.errstate(over="ignore"):
        f32 = np.float32(v)
    if float(f32) != v:
        return b"\xfb" + struct.pack(">d", v)
    with np.errstate(over="ignore"):
        f16 = np.float16(f32)
    if np.isfinite(f16) and np.float32(f16) == f32:
        return b"\xf9" + np.asarray(f16, dtype=">f2").tobytes()
    return b"\xfa" + struct.pack(">f", float(f32))


def _enc(x: Any, out: bytearray, depth: int) -> None:
    if depth > _MAX_DEPTH:
        raise CborError("value nesting too deep")
    if x is None:
        out += b"\xf6"
    elif x is True:
        out += b"\xf5"
    elif x is False:
        out += b"\xf4"
    elif isinstance(x, (int, np.integer)):
        v = int(x)
        out += _head(0, v) if v >= 0 else _head(1, -1 - v)
    elif isinstance(x, (float, np.floating)):
        out += _float(float(x))
    elif isinstance(x, str):
        raw = x.encode("utf-8")
        out += _head(3, len(raw)) + raw
    elif isinstance(x, (bytes, bytearray)):
        out += _head(2, len(x)) + bytes(x)
    elif isinstance(x, (list, tuple)):
        out += _head(4, len(x))
        for v in x:
            _enc(v, out, depth + 1)
    elif isinstance(x, dict):
        out += _head(5, len(x))
        for k, v in x.items():
            _enc(k, out, depth + 1)
            _enc(v, out, depth + 1)
    else:
        raise CborError(f"cannot encode {type(x).__name__}")


def cbor_dumps(x: Any) -> bytes:
    """Encode ``x`` (None, bool, int, float, str, bytes, list/tuple, dict)."""
    out = bytearray()
    _enc(x, out, 0)
    return bytes(out)


# ---------------------------------------------------------------- circuit ops
@dataclass(frozen=True)
class Mat:
    """``Uni::Mat(k, matrix)``: an explicit 2x2 (on qubit k) or 4x4 (on
    qubits k, k+1, qubit k the most significant bit) unitary."""
    k: int
    mat: np.ndarray

    def __eq__(self, other) -> bool:
        return (isinstance(other, Mat) and self.k == other.k
                and np.array_equal(np.asarray(self.mat), np.asarray(other.mat)))

    def __hash__(self) -> int:
        return hash((self.k, np.asarray(self.mat).tobytes()))


# gate.rs:25-59: variant name -> number of float parameters after the index
_GATE_PARAMS = {"U": 3, "H": 0, "X": 0, "Z": 0, "S": 0, "SInv": 0, "XRot": 1, "ZRot": 1,
                "CX": 0, "CXRev": 0, "CZ": 0, "Haar2": 0}
_GATE_Q2 = ("CX", "CXRev", "CZ", "Haar2")


@dataclass(frozen=True)
class Gate:
    """``Uni::Gate(Gate)``: a named gate on qubit k (two-qubit gates on k,
    k+1), with its angles (``U``: alpha, beta, gamma; ``XRot``/``ZRot``: the
    angle)."""
    name: str
    k: int
    params: Tuple[float, ...] = ()

    def __post_init__(self):
        if self.name not in _GATE_PARAMS:
            raise ValueError(f"unknown gate {self.name!r}")
        if len(self.params) != _GATE_PARAMS[self.name]:
            raise ValueError(f"gate {self.name} takes {_GATE_PARAMS[self.name]} angle(s)")

    @property
    def is_q2(self) -> bool:
        return self.name in _GATE_Q2

    def matrix(self, rng: Optional[np.random.Generator] = None) -> np.ndarray:
        return gate_matrix(self.name, self.params, rng)


@dataclass(frozen=True)
class Rand:
    """``Meas::Rand(k)``: a Born-sampled Z measurement of qubit k."""
    k: int


@dataclass(frozen=True)
class Proj:
    """``Meas::Proj(k, outcome)`` (the workspace's ``Meas::Postsel(k, bool)``):
    a projection of qubit k onto ``outcome`` (0 or 1)."""
    k: int
    outcome: int


Uni = Union[Mat, Gate]
Meas = Union[Rand, Proj]


def _cis(a: float) -> complex:
    return complex(math.cos(a), math.sin(a))


def gate_matrix(name: str, params: Sequence[float] = (), rng=None) -> np.ndarray:
    """``Gate::into_matrix_rng`` (gate.rs:226-265) with the constructors of
    gate.rs:606-855.  ``X`` is the identity, as ``make_x`` (gate.rs:649-658)
    writes it.  ``Haar2`` draws from ``rng`` with this package's generator
    (a QR of a complex Ginibre matrix with the phases of R's diagonal
    divided out, gate.rs:888-909), not from the reference's random stream."""
    if name == "U":
        alpha, beta, gamma = (float(x) for x in params)
        b2 = beta / 2.0
        pre = _cis(b2)
        on0 = complex(math.cos(b2))
        on1 = _cis(alpha + gamma) * on0
        off = complex(math.sin(b2))
        off0 = -1j * _cis(gamma) * off
        off1 = -1j * _cis(alpha) * off
        return np.array([[pre * on0, pre * off0], [pre * off1, pre * on1]], dtype=np.complex128)
    if name == "H":
        h = math.sqrt(1.0 / 2.0)
        return np.array([[h, h], [h, -h]], dtype=np.complex128)
    if name == "X":
        return np.eye(2, dtype=np.complex128)
    if name == "Z":
        return np.diag([1.0, -1.0]).astype(np.complex128)
    if name == "S":
        return np.diag([1.0, 1j]).astype(np.complex128)
    if name == "SInv":
        return np.diag([1.0, -1j]).astype(np.complex128)
    if name == "XRot":
        a2 = float(params[0]) / 2.0
        pre = _cis(a2)
        on = complex(math.cos(a2))
        off = -1j * complex(math.sin(a2))
        return np.array([[pre * on, pre * off], [pre * off, pre * on]], dtype=np.complex128)
    if name == "ZRot":
        return np.diag([1.0, _cis(float(params[0]))]).astype(np.complex128)
    if name == "CX":
        m = np.eye(4, dtype=np.complex128)
        m[2:, 2:] = [[0, 1], [1, 0]]
        return m
    if name == "CXRev":
        m = np.zeros((4, 4), dtype=np.complex128)
        m[0, 0] = m[1, 3] = m[2, 2] = m[3, 1] = 1.0
        return m
    if name == "CZ":
        return np.diag([1.0, 1.0, 1.0, -1.0]).astype(np.complex128)
    if name == "Haar2":
        from .workload import haar4
        if rng is None:
            raise ValueError("Haar2 needs a generator")
        return haar4(rng)
    raise ValueError(f"unknown gate {name!r}")


# ---------------------------------------------------------------- serde mapping
def _array2_to_serde(a: np.ndarray) -> dict:
    a = np.asarray(a, dtype=np.complex128)
    if a.ndim != 2:
        raise ValueError("expected a matrix")
    return {"v": 1, "dim": [int(a.shape[0]), int(a.shape[1])],
            "data": [[float(z.real), float(z.imag)] for z in a.reshape(-1)]}


def _array2_from_serde(v: Any) -> np.ndarray:
    if isinstance(v, (list, tuple)) and len(v) == 3:  # field order v, dim, data
        ver, dim, data = v
    elif isinstance(v, dict):
        ver, dim, data = v.get("v"), v.get("dim"), v.get("data")
    else:
        raise CborError("not an ndarray")
    if ver != 1:
        raise CborError(f"unsupported ndarray format version {ver!r}")
    if not isinstance(dim, (list, tuple)) or len(dim) != 2:
        raise CborError("expected a two-dimensional ndarray")
    r, c = int(dim[0]), int(dim[1])
    if not isinstance(data, (list, tuple)) or len(data) != r * c:
        raise CborError("ndarray data length does not match its shape")
    out = np.empty(r * c, dtype=np.complex128)
    for i, z in enumerate(data):
        if not isinstance(z, (list, tuple)) or len(z) != 2:
            raise CborError("expected a [re, im] pair")
        out[i] = complex(float(z[0]), float(z[1]))
    return out.reshape(r, c)


def _variant(v: Any) -> Tuple[str, Any]:
    """(name, payload) of an externally tagged enum value."""
    if isinstance(v, str):
        return v, None
    if isinstance(v, dict) and len(v) == 1:
        (name, payload), = v.items()
        if not isinstance(name, str):
            raise CborError("enum variant names must be text")
        return name, payload
    raise CborError("expected an externally tagged enum value")


def _outcome_from_serde(v: Any) -> int:
    if v in ("Zero", 0):
        return 0
    if v in ("One", 1):
        return 1
    raise CborError(f"bad Outcome {v!r}")


def uni_to_serde(u: Uni) -> dict:
    if isinstance(u, Mat):
        return {"Mat": [int(u.k), _array2_to_serde(u.mat)]}
    if isinstance(u, Gate):
        if u.params:
            return {"Gate": {u.name: [int(u.k), *[float(p) for p in u.params]]}}
        return {"Gate": {u.name: int(u.k)}}
    raise TypeError("expected Mat or Gate")


def uni_from_serde(v: Any) -> Uni:
    name, payload = _variant(v)
    if name == "Mat":
        if not isinstance(payload, (list, tuple)) or len(payload) != 2:
            raise CborError("Uni::Mat holds (usize, Array2)")
        return Mat(int(payload[0]), _array2_from_serde(payload[1]))
    if name == "Gate":
        gname, gp = _variant(payload)
        if gname not in _GATE_PARAMS:
            raise CborError(f"unknown gate {gname!r}")
        if _GATE_PARAMS[gname] == 0:
            return Gate(gname, int(gp))
        if not isinstance(gp, (list, tuple)) or len(gp) != 1 + _GATE_PARAMS[gname]:
            raise CborError(f"gate {gname} holds its index and {_GATE_PARAMS[gname]} angle(s)")
        return Gate(gname, int(gp[0]), tuple(float(x) for x in gp[1:]))
    raise CborError(f"unknown Uni variant {name!r}")


def meas_to_serde(m: Meas) -> dict:
    if isinstance(m, Rand):
        return {"Rand": int(m.k)}
    if isinstance(m, Proj):
        return {"Proj": [int(m.k), "One" if m.outcome else "Zero"]}
    raise TypeError("expected Rand or Proj")


def meas_from_serde(v: Any) -> Meas:
    name, payload = _variant(v)
    if name == "Rand":
        return Rand(int(payload))
    if name == "Proj":
        if not isinstance(payload, (list, tuple)) or len(payload) != 2:
            raise CborError("Meas::Proj holds (usize, Outcome)")
        return Proj(int(payload[0]), _outcome_from_serde(payload[1]))
    raise CborError(f"unknown Meas variant {name!r}")


# ---------------------------------------------------------------- cbor files
def save_cbor(obj: Any, path) -> None:
    """``save_cbor`` (circuit.rs:476-497): existing files are overwritten."""
    with open(path, "wb") as f:
        f.write(cbor_dumps(obj))


def load_cbor(path) -> Any:
    """``load_cbor`` (circuit.rs:499-517), as the serde data model."""
    with open(path, "rb") as f:
        return cbor_loads(f.read())


def _layers(v: Any, what: str) -> list:
    if not isinstance(v, list) or not all(isinstance(layer, list) for layer in v):
        raise CborError(f"expected a list of {what} layers")
    return v


def save_uni_circuit(path, layers: Sequence[Sequence[Uni]]) -> None:
    save_cbor([[uni_to_serde(u) for u in layer] for layer in layers], path)


def load_uni_circuit(path) -> List[List[Uni]]:
    """A ``Vec<Vec<Uni>>`` file (``MiptManifest::save_all``, haar.rs:277-294)."""
    return [[uni_from_serde(u) for u in layer] for layer in _layers(load_cbor(path), "unitary")]


def save_meas_layers(path, layers: Sequence[Sequence[Meas]]) -> None:
    save_cbor([[meas_to_serde(m) for m in layer] for layer in layers], path)


def load_meas_layers(path) -> List[List[Meas]]:
    """A ``Vec<Vec<Meas>>`` file (haar.rs:285-291)."""
    return [[meas_from_serde(m) for m in layer] for layer in _layers(load_cbor(path), "measurement")]


@dataclass
class MiptManifest:
    """A circuit set: ``MiptManifest`` (workspace/lib/haar.rs:211-320).

    Loading and saving keep every field.  The reference derives ``id`` from a
    hash of the other fields through Rust's ``DefaultHasher`` and ``StdRng``;
    here it is given."""
    seed: int
    nqubits: int
    depth: int
    num_circs: int
    p_meas: List[float]
    id: str

    def unis_fname(self, circ: int) -> str:
        return f"unis_seed={self.seed}_nqubits={self.nqubits}_depth={self.depth}_circ={circ}.cbor"

    def meas_fname(self, p: float, circ: int) -> str:
        return (f"meas_seed={self.seed}_nqubits={self.nqubits}_depth={self.depth}"
                f"_p={p:.6f}_circ={circ}.cbor")

    @staticmethod
    def manifest_fmt(seed: int, nqubits: int, depth: int, id_: str) -> str:
        return f"manifest_seed={seed}_nqubits={nqubits}_depth={depth}_id={id_}.cbor"

    def manifest_fname(self) -> str:
        return self.manifest_fmt(self.seed, self.nqubits, self.depth, self.id)

    def to_serde(self) -> dict:
        return {"seed": int(self.seed), "nqubits": int(self.nqubits), "depth": int(self.depth),
                "num_circs": int(self.num_circs), "p_meas": [float(p) for p in self.p_meas],
                "id": str(self.id)}

    @classmethod
    def from_serde(cls, v: Any) -> "MiptManifest":
        if isinstance(v, (list, tuple)) and len(v) == 6:
            v = dict(zip(("seed", "nqubits", "depth", "num_circs", "p_meas", "id"), v))
        if not isinstance(v, dict):
            raise CborError("expected a MiptManifest map")
        try:
            return cls(int(v["seed"]), int(v["nqubits"]), int(v["depth"]), int(v["num_circs"]),
                       [float(p) for p in v["p_meas"]], str(v["id"]))
        except KeyError as e:
            raise CborError(f"MiptManifest lacks field {e.args[0]!r}") from None

    def save(self, outdir) -> str:
        path = os.path.join(outdir, self.manifest_fname())
        save_cbor(self.to_serde(), path)
        return path

    @classmethod
    def load(cls, path) -> "MiptManifest":
        return cls.from_serde(load_cbor(path))

    @classmethod
    def load_in(cls, seed: int, nqubits: int, depth: int, id_: str, indir) -> "MiptManifest":
        return cls.load(os.path.join(indir, cls.manifest_fmt(seed, nqubits, depth, id_)))

    def save_all(self, outdir, rng: np.random.Generator) -> None:
        """Write the set's circuits and measurement layers in the layout of
        ``MiptManifest::save_all`` (haar.rs:277-298): per circuit a
        brickwork of Haar unitaries (layer t offset when t is odd), then per
        p a layer list of ``Meas::Rand`` at each qubit with probability p,
        then the manifest.  The random stream is this package's, so the
        circuits differ from the reference's for the same seed."""
        from .workload import brickwork_sites, haar4
        for circ in range(self.num_circs):
            unis = [[Mat(k, haar4(rng)) for k in brickwork_sites(self.nqubits, t % 2 == 1)]
                    for t in range(self.depth)]
            save_uni_circuit(os.path.join(outdir, self.unis_fname(circ)), unis)
            for p in self.p_meas:
                meas = [[Rand(k) for k in range(self.nqubits) if rng.random() < p]
                        for _ in range(self.depth)]
                save_meas_layers(os.path.join(outdir, self.meas_fname(p, circ)), meas)
        self.save(outdir)


# ---------------------------------------------------------------- npz files
def save_unis_npz(path, layers: Sequence[Sequence[Tuple[int, np.ndarray]]]) -> None:
    """``save_unis`` (haar.rs:54-77): ``depth`` (u32), ``targs`` (u32) and
    ``mats`` (complex128, count x d x d), one entry per unitary."""
    depth, targs, mats = [], [], []
    for t, layer in enumerate(layers):
        for k, m in layer:
            depth.append(t)
            targs.append(int(k))
            mats.append(np.asarray(m, dtype=np.complex128))
    if not mats:
        raise ValueError("no unitaries to stack")
    np.savez(path, depth=np.asarray(depth, dtype=np.uint32),
             targs=np.asarray(targs, dtype=np.uint32), mats=np.stack(mats))


def load_unis_npz(path) -> List[List[Tuple[int, np.ndarray]]]:
    """``load_unis`` (haar.rs:79-100): consecutive entries of equal depth
    form a layer, so a layer without unitaries does not come back."""
    with np.load(path, allow_pickle=False) as z:
        depth, targs, mats = z["depth"], z["targs"], z["mats"]
    layers: List[List[Tuple[int, np.ndarray]]] = []
    prev = None
    for d, k, m in zip(depth.tolist(), targs.tolist(), mats):
        if prev is None or d != prev:
            layers.append([])
            prev = d
        layers[-1].append((int(k), np.array(m, dtype=np.complex128)))
    return layers


def save_meas_npz(path, per_p: Iterable[Tuple[float, Sequence[Sequence[Meas]]]]) -> None:
    """``save_meas`` (workspace/lib/lib.rs:61-103): per layer its p
    (``pvals``, f64) and length (``layer_lens``, u32); per measurement its
    qubit (``targs``, u32) and ``ops`` (i32: -1 sampled, 0/1 post-selected)."""
    pvals, lens, targs, ops = [], [], [], []
    for p, layers in per_p:
        for layer in layers:
            for m in layer:
                targs.append(int(m.k))
                ops.append(-1 if isinstance(m, Rand) else int(bool(m.outcome)))
            pvals.append(float(p))
            lens.append(len(layer))
    np.savez(path, pvals=np.asarray(pvals, dtype=np.float64),
             layer_lens=np.asarray(lens, dtype=np.uint32),
             targs=np.asarray(targs, dtype=np.uint32), ops=np.asarray(ops, dtype=np.int32))


def load_meas_npz(path) -> Tuple[np.ndarray, List[List[List[Meas]]]]:
    """``load_meas`` (workspace/lib/lib.rs:105-147): consecutive layers of
    equal p form one p's layer list; ``ops`` 0/1 is a post-selection, any
    other value a sampled measurement."""
    with np.load(path, allow_pickle=False) as z:
        pvals, lens, targs, ops = z["pvals"], z["layer_lens"], z["targs"], z["ops"]
    ps: List[float] = []
    out: List[List[List[Meas]]] = []
    it = iter(zip(targs.tolist(), ops.tolist()))
    prev = None
    for p, ln in zip(pvals.tolist(), lens.tolist()):
        if prev is None or p != prev:
            ps.append(p)
            out.append([])
            prev = p
        layer: List[Meas] = []
        for _ in range(int(ln)):
            try:
                k, op = next(it)
            except StopIteration:
                raise ValueError("operations list ended unexpectedly") from None
            layer.append(Proj(int(k), int(op)) if op in (0, 1) else Rand(int(k)))
        out[-1].append(layer)
    return np.asarray(ps, dtype=np.float64), out
