"""Circuit and record files (tnmps.formats), on the CPU.

The CBOR codec is pinned by the examples of RFC 8949 Appendix A (the
encoder writes the preferred, shortest-exact serialization, which is also
serde_cbor 0.11's float rule).  The serde layouts are pinned by byte strings
written out by hand from serde's data model: externally tagged enums,
structs as maps, ndarray as {v, dim, data}, Complex64 as [re, im].  The npz
layouts follow the reference's own round-trip test
(workspace/src/circuit_file_test.rs:23-44).
"""
import math

import numpy as np
import pytest

from tnmps import formats as F
from tnmps.workload import haar4

# RFC 8949 Appendix A: (value, encoding) pairs the encoder must reproduce
RFC_ENCODE = [
    (0, "00"), (1, "01"), (10, "0a"), (23, "17"), (24, "1818"), (25, "1819"), (100, "1864"),
    (1000, "1903e8"), (1000000, "1a000f4240"), (1000000000000, "1b000000e8d4a51000"),
    (18446744073709551615, "1bffffffffffffffff"), (-1, "20"), (-10, "29"), (-100, "3863"),
    (-1000, "3903e7"), (0.0, "f90000"), (-0.0, "f98000"), (1.0, "f93c00"),
    (1.1, "fb3ff199999999999a"), (1.5, "f93e00"), (65504.0, "f97bff"),
    (100000.0, "fa47c35000"), (3.4028234663852886e+38, "fa7f7fffff"),
    (1.0e+300, "fb7e37e43c8800759c"), (5.960464477539063e-8, "f90001"),
    (0.00006103515625, "f90400"), (-4.0, "f9c400"), (-4.1, "fbc010666666666666"),
    (math.inf, "f97c00"), (-math.inf, "f9fc00"), (False, "f4"), (True, "f5"), (None, "f6"),
    ("", "60"), ("a", "6161"), ("IETF", "6449455446"), ('"\\', "62225c"),
    ("ü", "62c3bc"), ("水", "63e6b0b4"), (b"", "40"), (b"\x01\x02\x03\x04", "4401020304"),
    ([], "80"), ([1, 2, 3], "83010203"), ([1, [2, 3], [4, 5]], "8301820203820405"),
    (list(range(1, 26)), "98190102030405060708090a0b0c0d0e0f101112131415161718181819"),
    ({}, "a0"), ({1: 2, 3: 4}, "a201020304"), ({"a": 1, "b": [2, 3]}, "a26161016162820203"),
    (["a", {"b": "c"}], "826161a161626163"),
    ({"a": "A", "b": "B", "c": "C", "d": "D", "e": "E"},
     "a56161614161626142616361436164614461656145"),
]

# RFC 8949 Appendix A: decode-only forms (indefinite lengths, tags, simple values)
RFC_DECODE = [
    ("5f42010243030405ff", b"\x01\x02\x03\x04\x05"),
    ("7f657374726561646d696e67ff", "streaming"),
    ("9fff", []), ("9f018202039f0405ffff", [1, [2, 3], [4, 5]]),
    ("9f01820203820405ff", [1, [2, 3], [4, 5]]), ("83018202039f0405ff", [1, [2, 3], [4, 5]]),
    ("bf61610161629f0203ffff", {"a": 1, "b": [2, 3]}),
    ("826161bf61626163ff", ["a", {"b": "c"}]),
    ("bf6346756ef563416d7421ff", {"Fun": True, "Amt": -2}),
    ("c249010000000000000000", 18446744073709551616),
    ("3bffffffffffffffff", -18446744073709551616),
    ("c349010000000000000000", -18446744073709551617),
    ("c11a514b67b0", 1363896240), ("d74401020304", b"\x01\x02\x03\x04"),
    ("fa7f800000", math.inf), ("fb7ff0000000000000", math.inf), ("f97bff", 65504.0),
    ("fb3ff199999999999a", 1.1), ("f90001", 5.960464477539063e-8),
]


@pytest.mark.parametrize("value,hexenc", RFC_ENCODE)
def test_cbor_rfc8949_encode(value, hexenc):
    assert F.cbor_dumps(value).hex() == hexenc


@pytest.mark.parametrize("value,hexenc", RFC_ENCODE)
def test_cbor_rfc8949_decode(value, hexenc):
    got = F.cbor_loads(bytes.fromhex(hexenc))
    if isinstance(value, float):
        assert got == value and math.copysign(1, got) == math.copysign(1, value)
    else:
        assert got == value


@pytest.mark.parametrize("hexenc,value", RFC_DECODE)
def test_cbor_rfc8949_decode_only(hexenc, value):
    assert F.cbor_loads(bytes.fromhex(hexenc)) == value


def test_cbor_nan_and_errors():
    assert F.cbor_dumps(math.nan).hex() == "f97e00"
    assert math.isnan(F.cbor_loads(bytes.fromhex("f97e00")))
    assert math.isnan(F.cbor_loads(bytes.fromhex("fb7ff8000000000000")))
    for bad in ("", "1b0000", "83 01 02", "ff", "8301", "a16161", "01 02", "7f61ff"):
        with pytest.raises(F.CborError):
            F.cbor_loads(bytes.fromhex(bad.replace(" ", "")))


def test_serde_meas_layout():
    """Meas::Rand(3) = {"Rand": 3}; Meas::Proj(5, One) = {"Proj": [5, "One"]}."""
    assert F.cbor_dumps(F.meas_to_serde(F.Rand(3))).hex() == "a16452616e6403"
    assert F.cbor_dumps(F.meas_to_serde(F.Proj(5, 1))).hex() == "a16450726f6a8205634f6e65"
    assert F.cbor_dumps(F.meas_to_serde(F.Proj(0, 0))).hex() == "a16450726f6a8200645a65726f"
    assert F.meas_from_serde(F.cbor_loads(bytes.fromhex("a16452616e6403"))) == F.Rand(3)
    # serde also accepts a variant by index and an Outcome by its discriminant
    assert F.meas_from_serde({"Proj": [2, 1]}) == F.Proj(2, 1)


def test_serde_mat_layout():
    """Uni::Mat(1, I2) = {"Mat": [1, {"v": 1, "dim": [2, 2], "data": [[1,0],[0,0],[0,0],[1,0]]}]}."""
    enc = F.cbor_dumps(F.uni_to_serde(F.Mat(1, np.eye(2))))
    one, zero = "82f93c00f90000", "82f90000f90000"
    want = ("a1634d6174" "82" "01" "a3" "6176" "01" "6364696d" "820202" "6464617461" "84"
            + one + zero + zero + one)
    assert enc.hex() == want
    back = F.uni_from_serde(F.cbor_loads(enc))
    assert back == F.Mat(1, np.eye(2))


def test_serde_gate_layout():
    """Uni::Gate(Gate::CX(4)) = {"Gate": {"CX": 4}};
    Uni::Gate(Gate::XRot(0, 0.5)) = {"Gate": {"XRot": [0, 0.5]}}."""
    assert F.cbor_dumps(F.uni_to_serde(F.Gate("CX", 4))).hex() == "a16447617465a162435804"
    assert (F.cbor_dumps(F.uni_to_serde(F.Gate("XRot", 0, (0.5,)))).hex()
            == "a16447617465a16458526f748200f93800")
    for g in (F.Gate("U", 2, (0.1, 0.2, 0.3)), F.Gate("Haar2", 7), F.Gate("SInv", 1)):
        assert F.uni_from_serde(F.cbor_loads(F.cbor_dumps(F.uni_to_serde(g)))) == g
    with pytest.raises(ValueError):
        F.Gate("XRot", 0)


def test_gate_matrices():
    """gate.rs:606-855: unitarity, the X quirk (identity, gate.rs:649-658),
    U(alpha, beta, gamma) equal to the matrix product ZRot(alpha) XRot(beta)
    ZRot(gamma) (the operator applies ZRot(gamma) first, as gate.rs:603-605
    says), and the basis convention (qubit k the most significant bit)."""
    rng = np.random.default_rng(3)
    for name, ps in (("U", (0.3, 1.1, -0.7)), ("H", ()), ("X", ()), ("Z", ()), ("S", ()),
                     ("SInv", ()), ("XRot", (0.9,)), ("ZRot", (2.1,)), ("CX", ()),
                     ("CXRev", ()), ("CZ", ()), ("Haar2", ())):
        m = F.gate_matrix(name, ps, rng)
        assert np.allclose(m.conj().T @ m, np.eye(len(m)), atol=1e-14)
    assert np.array_equal(F.gate_matrix("X"), np.eye(2))
    a, b, c = 0.3, 1.1, -0.7
    u = F.gate_matrix("U", (a, b, c))
    zr = lambda t: F.gate_matrix("ZRot", (t,))
    xr = F.gate_matrix("XRot", (b,))
    assert np.allclose(u, zr(a) @ xr @ zr(c), atol=1e-15)
    assert not np.allclose(u, zr(c) @ xr @ zr(a), atol=1e-3)
    # XRot(pi) is X up to a phase
    assert np.allclose(np.abs(F.gate_matrix("XRot", (math.pi,))), [[0, 1], [1, 0]], atol=1e-15)
    cx = F.gate_matrix("CX")
    assert cx[3, 2] == 1 and cx[2, 3] == 1       # |10> <-> |11>: control is qubit k
    cxr = F.gate_matrix("CXRev")
    assert cxr[1, 3] == 1 and cxr[3, 1] == 1     # |01> <-> |11>: control is qubit k+1
    assert np.allclose(F.gate_matrix("H") @ F.gate_matrix("H"), np.eye(2), atol=1e-15)


def test_cbor_circuit_roundtrip(tmp_path):
    rng = np.random.default_rng(5)
    unis = [[F.Mat(k, haar4(rng)) for k in range(t % 2, 5, 2)] for t in range(4)]
    unis[2].append(F.Gate("H", 0))
    meas = [[F.Rand(1), F.Rand(4)], [], [F.Proj(0, 1), F.Proj(3, 0)], [F.Rand(2)]]
    F.save_uni_circuit(tmp_path / "u.cbor", unis)
    F.save_meas_layers(tmp_path / "m.cbor", meas)
    assert F.load_uni_circuit(tmp_path / "u.cbor") == unis
    assert F.load_meas_layers(tmp_path / "m.cbor") == meas


def test_manifest(tmp_path):
    m = F.MiptManifest(10546, 8, 6, 2, [0.1, 0.25], "00c0ffee12345678")
    assert m.unis_fname(1) == "unis_seed=10546_nqubits=8_depth=6_circ=1.cbor"
    assert m.meas_fname(0.1, 0) == "meas_seed=10546_nqubits=8_depth=6_p=0.100000_circ=0.cbor"
    assert m.manifest_fname() == "manifest_seed=10546_nqubits=8_depth=6_id=00c0ffee12345678.cbor"
    m.save_all(tmp_path, np.random.default_rng(0))
    back = F.MiptManifest.load_in(10546, 8, 6, "00c0ffee12345678", tmp_path)
    assert back == m
    # struct field order of haar.rs:211-218
    assert list(F.load_cbor(tmp_path / m.manifest_fname()).keys()) == \
        ["seed", "nqubits", "depth", "num_circs", "p_meas", "id"]
    unis = F.load_uni_circuit(tmp_path / m.unis_fname(1))
    assert [[u.k for u in layer] for layer in unis] == [[0, 2, 4, 6], [1, 3, 5]] * 3
    meas = F.load_meas_layers(tmp_path / m.meas_fname(0.25, 1))
    assert len(meas) == 6 and all(isinstance(x, F.Rand) for layer in meas for x in layer)
    assert all([x.k for x in layer] == sorted(x.k for x in layer) for layer in meas)


def test_unis_npz_roundtrip(tmp_path):
    """circuit_file_test.rs:26-31: save then load gives the same layers; the
    array dtypes are the reference's (u32, u32, complex128)."""
    rng = np.random.default_rng(7)
    n, T = 6, 12
    layers = [[(k, haar4(rng)) for k in range(t % 2, n - 1, 2)] for t in range(T)]
    F.save_unis_npz(tmp_path / "unis.npz", layers)
    back = F.load_unis_npz(tmp_path / "unis.npz")
    assert len(back) == T
    for a, b in zip(layers, back):
        assert [k for k, _ in a] == [k for k, _ in b]
        assert all(np.array_equal(x, y) for (_, x), (_, y) in zip(a, b))
    with np.load(tmp_path / "unis.npz") as z:
        assert z["depth"].dtype == np.uint32 and z["targs"].dtype == np.uint32
        assert z["mats"].dtype == np.complex128 and z["mats"].shape == (30, 4, 4)
    # a layer without unitaries does not come back (chunk_by on depth)
    F.save_unis_npz(tmp_path / "gap.npz", [layers[0], [], layers[2]])
    assert len(F.load_unis_npz(tmp_path / "gap.npz")) == 2


def test_meas_npz_roundtrip(tmp_path):
    """circuit_file_test.rs:33-43 with post-selections mixed in; equal
    consecutive p values merge (chunk_by on pvals)."""
    rng = np.random.default_rng(8)
    ps = [0.01, 0.1, 0.25]
    per_p = []
    for p in ps:
        layers = []
        for _ in range(5):
            layer = []
            for k in range(6):
                if rng.random() < 0.5:
                    layer.append(F.Rand(k) if rng.random() < 0.5 else F.Proj(k, int(rng.random() < 0.5)))
            layers.append(layer)
        per_p.append((p, layers))
    F.save_meas_npz(tmp_path / "meas.npz", per_p)
    p_read, back = F.load_meas_npz(tmp_path / "meas.npz")
    assert list(p_read) == ps
    assert back == [layers for _, layers in per_p]
    with np.load(tmp_path / "meas.npz") as z:
        assert z["ops"].dtype == np.int32 and z["pvals"].dtype == np.float64
        assert set(z["ops"].tolist()) <= {-1, 0, 1}
    F.save_meas_npz(tmp_path / "merge.npz", [(0.2, [[F.Rand(0)]]), (0.2, [[F.Rand(1)]])])
    p2, b2 = F.load_meas_npz(tmp_path / "merge.npz")
    assert list(p2) == [0.2] and b2 == [[[F.Rand(0)], [F.Rand(1)]]]
