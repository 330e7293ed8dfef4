"""Host logic of the serving loop (no device needed)."""
from tnmps.serve import layer_cost, may_admit, mps_bytes


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
