"""Host logic of the launch-shape autotuner (no device needed)."""
import json

import pytest

from tnmps import _lib as L
from tnmps.autotune import CANDIDATES, bucket_of, load_table


def _rows():
    return [dict(c_max=c, theta_tile=32, bw=8, split_max=8, split_rows=128, ahead=2, qr=1)
            for c in (16, 32, 64, 128, 256, 512, 1024)]


def test_bucket_of():
    t = _rows()
    assert bucket_of(t, 1) == 0
    assert bucket_of(t, 16) == 0
    assert bucket_of(t, 17) == 1
    assert bucket_of(t, 512) == 5
    assert bucket_of(t, 513) == 6
    assert bucket_of(t, 5000) == 6  # the last row takes larger C


def test_candidates_are_valid_fields():
    assert set(CANDIDATES) <= set(L.TUNE_FIELDS)
    assert all(v in (32, 64) for v in CANDIDATES["theta_tile"])
    assert all(v in (8, 16) for v in CANDIDATES["bw"])
    assert all(1 <= v <= 8 for v in CANDIDATES["split_max"])
    assert all(v >= 64 for v in CANDIDATES["split_rows"])
    assert all(1 <= v <= 3 for v in CANDIDATES["ahead"])
    assert set(CANDIDATES["qr"]) <= {0, 1}


def test_load_table(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"k": {"table": _rows()}}))
    assert load_table(str(p), "k") == _rows()
    # a row without a field that has a default (an older table) takes it
    old = [{f: v for f, v in r.items() if f != "qr"} for r in _rows()]
    p.write_text(json.dumps({"k": {"table": old}}))
    assert load_table(str(p), "k") == _rows()
    bad = _rows()
    bad[0]["extra"] = 1
    p.write_text(json.dumps({"k": {"table": bad}}))
    with pytest.raises(ValueError):
        load_table(str(p), "k")
