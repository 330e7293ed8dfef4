"""Host logic of the serving loop (no device needed)."""
from tnmps.serve import layer_cost, may_admit, mps_bytes, workspace_bytes


def test_may_admit_least_loaded():
    assert may_admit([0.0, 5.0], 0)
    assert not may_admit([0.0, 5.0], 1)
    assert may_admit([3.0, 3.0], 0) and may_admit([3.0, 3.0], 1)  # ties admit
    assert may_admit([7.0], 0)


def test_layer_cost_grows_with_reachable_bond():
    # chi_eff = min(chi, 2^(n/2)): a 16-qubit chain never exceeds 256
    assert layer_cost(16, 256) == layer_cost(16, 512) == 16 * 256.0 ** 3
    assert layer_cost(64, 256) == 64 * 256.0 ** 3
    assert layer_cost(8, 64) == 8 * 16.0 ** 3


def test_mps_bytes_product_chain():
    # chi = 1: every Gamma is 1 x 2 x 1 (32 B), every lambda one double
    n = 5
    assert mps_bytes(n, 1) == 16 * 2 * n + 8 * (n - 1) + 4 * (n + 1)


def test_workspace_bytes():
    # chi = 256: a slot is 2 x (512 x 512) complex plus per-column arrays
    per = 16.0 * 2 * 512 * 512 + 64.0 * 512 + 4096.0
    gp = (512 // 16 // 2 + 1) * 8 * 4 * 256 * 16.0  # the wider panels' partial Grams
    assert workspace_bytes(256, 10 * per) == 10 * per + 10 * gp
    assert workspace_bytes(256, 1000 * per) == 1000 * per + 256 * gp
    assert workspace_bytes(256, 1.0) == per + gp  # at least one slot
