"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/tnmps.h declares, and rejects work without a device.  No compute
call is made here."""
import re
from pathlib import Path

import pytest

from tnmps import _lib
from tnmps.mps import MPS, MPSError

HEADER = Path(__file__).resolve().parent.parent / "include" / "tnmps.h"


def declared_functions():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tn_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("tn_mps_new_qubits", "tn_apply_u2", "tn_meas_layer", "tn_entropy_vn",
                 "tn_run_circuit", "tn_schmidt"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} not exported by libtnmps.so"


def test_python_binding_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_status_strings():
    assert _lib.status_str(0) == "ok"
    assert "zero probability" in _lib.status_str(_lib.TN_ZERO_PROB)
    assert _lib.status_str(999) == "unknown status"


def test_empty_system_is_rejected_before_device():
    """MPS::new (mps.rs:375): n == 0 is EmptySystem."""
    with pytest.raises(MPSError) as ei:
        MPS.new_qubits(0, ("const", 4))
    assert ei.value.kind == "EmptySystem"


def test_no_device_fails_loudly():
    lib = _lib.load()
    if lib.tn_device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(MPSError) as ei:
        MPS.new_qubits(4, ("const", 4))
    assert ei.value.kind == "Device"
