"""Co-evolution drivers and reference-format circuits on the GPU, against the
oracle.  The oracle side is restated here from the reference drivers
(workspace/src/haar_coev_probs.rs:33-83, haar_coev_probs_batching.rs:43-70,
workspace/lib/haar.rs:136-161), drawing the same uniforms in the same order.
"""
import numpy as np
import pytest

from oracle.mps_oracle import VidalMPS, run_trajectory
from tnmps import coev
from tnmps import formats as F
from tnmps.workload import haar4

pytestmark = pytest.mark.gpu


def _close(got, want, rtol=1e-10, atol=1e-12):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape
    err = np.abs(got - want) - (atol + rtol * np.abs(want))
    assert np.all(err <= 0), f"max excess {err.max():.3e}"


def _circuit(n, depth, p, seed):
    rng = np.random.default_rng(seed)
    unis = [[F.Mat(k, haar4(rng)) for k in range(t % 2, n - 1, 2)] for t in range(depth)]
    meas = [[F.Rand(k) for k in range(n) if rng.random() < p] for _ in range(depth)]
    return unis, meas


@pytest.mark.parametrize("n,depth,p,seed", [(8, 8, 0.3, 1), (10, 6, 0.2, 2)])
def test_coev_probs_matches_oracle(gpu, n, depth, p, seed):
    unis, meas = _circuit(n, depth, p, seed)
    runs, bonds = 3, (2, 4, None, 6)
    got = coev.coev_probs(n, unis, meas, runs=runs, bonds=bonds,
                          rng=np.random.default_rng(100 + seed))
    # the driver draws, per run in order, num_meas + n * depth uniforms
    num_meas = sum(len(layer) for layer in meas)
    rng = np.random.default_rng(100 + seed)
    layers = [([(u.k, u.mat) for u in unis[t]], [m.k for m in meas[t]]) for t in range(depth)]
    assert got["meas_locs"].tolist() == [[t, m.k] for t in range(depth) for m in meas[t]]
    for r in range(runs):
        draws = rng.random(num_meas + n * depth)
        ex = run_trajectory(n, layers, None, rs=draws)
        outs = [o for layer in ex["record"] for _, o in layer]
        assert got["traj_data"][r].tolist() == outs
        for b, chi in enumerate(bonds):
            if chi is None:
                want = [pr for layer in ex["probs"] for _, pr in layer]
            else:
                rep = run_trajectory(n, layers, ("const", chi), forced=ex["record"])
                want = [pr for layer in rep["probs"] for _, pr in layer]
            _close(got["prob_data"][r, b], want)


def test_run_manifest_files(gpu, tmp_path):
    """A circuit set written in the reference's layout (manifest + cbor
    files) replays through ``run_manifest``; the output file carries the
    reference's fields."""
    man = F.MiptManifest(10546, 8, 6, 2, [0.2, 0.3], "0123456789abcdef")
    man.save_all(tmp_path, np.random.default_rng(4))
    mfile = tmp_path / man.manifest_fname()
    m2, res = coev.run_manifest(mfile, 1, 0.3, runs=2, bonds=(2, None),
                                rng=np.random.default_rng(9))
    assert m2 == man
    unis = F.load_uni_circuit(tmp_path / man.unis_fname(1))
    meas = F.load_meas_layers(tmp_path / man.meas_fname(0.3, 1))
    again = coev.coev_probs(8, unis, meas, runs=2, bonds=(2, None),
                            rng=np.random.default_rng(9))
    for key in ("meas_locs", "traj_data", "prob_data"):
        assert np.array_equal(res[key], again[key])
    path = coev.save_coev_probs(tmp_path, man, str(mfile), 0.3, 1, "feedfacecafebeef", res)
    with np.load(path) as z:
        assert "".join(chr(c) for c in z["genid"]) == man.id
        assert z["prob_data"].shape == (2, 2, res["meas_locs"].shape[0])
        assert z["traj_data"].dtype == np.uint8 and z["meas_locs"].dtype == np.uint8
    with pytest.raises(ValueError):
        coev.run_manifest(mfile, 2, 0.3)
    with pytest.raises(ValueError):
        coev.run_manifest(mfile, 0, 0.25)


def test_coev_probs_batching_matches_oracle(gpu):
    n, depth, runs = 6, 5, 2
    rng = np.random.default_rng(21)
    unis = [[(k, haar4(rng)) for k in range(t % 2, n - 1, 2)] for t in range(depth)]
    meas_per_p = []
    for p in (0.2, 0.4):
        layers = [[F.Rand(k) for k in range(n) if rng.random() < p] for _ in range(depth)]
        free = min(set(range(n)) - {m.k for m in layers[1]})
        layers[1] = layers[1] + [F.Proj(free, 0)]  # a post-selection in the list
        meas_per_p.append(layers)
    bonds = (None, 2, 4)
    got = coev.coev_probs_batching(n, unis, meas_per_p, runs=runs, bonds=bonds,
                                   rng=np.random.default_rng(5))
    assert got["chi"].tolist() == [0, 2, 4]
    # oracle: apply_main_layer_probs (haar.rs:136-161), per measurement in
    # list order; one uniform per sampled measurement, in order
    drng = np.random.default_rng(5)
    for j, meas in enumerate(meas_per_p):
        for r in range(runs):
            st = VidalMPS(n, None)
            traj = np.zeros((depth, n), dtype=np.int8)
            prob = np.full((3, depth, n), -1.0)
            rec = []
            for t in range(depth):
                for k, m in unis[t]:
                    st.apply_unitary2(k, m)
                outs = []
                for m in meas[t]:
                    if isinstance(m, F.Rand):
                        probs = st.probs(m.k)
                        o, _ = st.measure_prob(m.k, drng.random())
                        pr = probs[o]
                    else:
                        o = m.outcome
                        pr = st.prob(m.k, o)
                        st.measure_postsel_prob(m.k, o)
                    outs.append(F.Proj(m.k, o))
                    traj[t, m.k] = 1 if o else -1
                    prob[0, t, m.k] = pr
                rec.append(outs)
            for b, chi in ((1, 2), (2, 4)):
                st = VidalMPS(n, ("const", chi))
                for t in range(depth):
                    for k, m in unis[t]:
                        st.apply_unitary2(k, m)
                    for m in rec[t]:
                        prob[b, t, m.k] = st.prob(m.k, m.outcome)
                        st.measure_postsel_prob(m.k, m.outcome)
            assert np.array_equal(got["traj"][j, r], traj)
            _close(got["prob"][j, r], prob)
