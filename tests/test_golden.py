"""Replay of the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle and checked there against the
dense brute force) through the oracle (CPU) and the HIP library (GPU)."""
from pathlib import Path

import numpy as np
import pytest

from helpers import assert_spectrum_close, state_overlap_err
from oracle.mps_oracle import VidalMPS, apply_bilayer

GOLD = Path(__file__).resolve().parent / "golden"


def load(name):
    return dict(np.load(GOLD / name, allow_pickle=False))


def unpack(f):
    n, depth = int(f["n"]), int(f["depth"])
    layers, forced = [], []
    uo = mo = 0
    for t in range(depth):
        cu = int(f["nu"][t])
        unis = [(int(f["uks"][uo + i]), f["mats"][uo + i]) for i in range(cu)]
        uo += cu
        cm = int(f["nm"][t])
        rec = [(int(f["mks"][mo + i]), int(f["outs"][mo + i])) for i in range(cm)]
        mo += cm
        layers.append((unis, [k for k, _ in rec]))
        forced.append(rec)
    return n, layers, forced


def check_traj(f, entropy, probs, spectra_fn, dims):
    np.testing.assert_allclose(entropy, f["entropy"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(probs, f["probs"], rtol=1e-10, atol=1e-12)
    assert list(dims) == list(f["dims"])
    for b, d in enumerate(f["dims"]):
        assert_spectrum_close(spectra_fn(b), f["spectra"][b, :d], msg=f"bond {b}")


@pytest.mark.parametrize("name", ["traj_n8_chi16.npz", "traj_n12_chi4.npz"])
def test_oracle_replays_fixture(name):
    f = load(name)
    n, layers, forced = unpack(f)
    st = VidalMPS(n, ("const", int(f["chi"])))
    ent = [st.entropy_vn(n // 2 - 1)]
    probs = []
    for (unis, _), rec in zip(layers, forced):
        res = apply_bilayer(st, unis, [(k, o, 0.0) for k, o in rec])
        probs += [x for _, _, x in res]
        ent.append(st.entropy_vn(n // 2 - 1))
    check_traj(f, ent, probs, lambda b: st.L[b], st.bond_dims())


def _gate_list(f):
    out = []
    for k, sz, m in zip(f["ks"], f["sizes"], f["mats"]):
        out.append((int(k), m[:2, :2] if sz == 2 else m))
    return out


def test_oracle_gates_fixture():
    f = load("gates_n6.npz")
    n = int(f["n"])
    st = VidalMPS(n, None)
    for k, m in _gate_list(f):
        if m.shape == (2, 2):
            st.apply_unitary1(k, m)
        else:
            st.apply_unitary2(k, m)
    assert state_overlap_err(st.contract(), f["psi"]) < 1e-13
    for b, d in enumerate(f["dims"]):
        assert_spectrum_close(st.L[b][:d], f["spectra"][b, :d])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["traj_n8_chi16.npz", "traj_n12_chi4.npz"])
def test_gpu_replays_fixture(gpu, name):
    from tnmps import MPS

    f = load(name)
    n, layers, forced = unpack(f)
    st = MPS.new_qubits(n, ("const", int(f["chi"])))
    res = st.run_circuit(layers, forced=forced)
    probs = [x for layer in res["probs"] for _, x in layer]
    check_traj(f, res["entropy"], probs, st.schmidt, st.bond_dims())


@pytest.mark.gpu
def test_gpu_gates_fixture(gpu):
    from tnmps import MPS

    f = load("gates_n6.npz")
    n = int(f["n"])
    st = MPS.new_qubits(n, None)
    st.apply_unis(_gate_list(f))
    assert state_overlap_err(st.contract(), f["psi"]) < 1e-12
    for b, d in enumerate(f["dims"]):
        assert_spectrum_close(st.schmidt(b)[:d], f["spectra"][b, :d])
