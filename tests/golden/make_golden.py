"""Write the golden fixtures under tests/golden/.

The reference's own tests hold no numeric vectors for this path (and its
Rust crate cannot be built here), so the fixtures come from the CPU oracle
(oracle/mps_oracle.py), each cross-checked against the dense state-vector
brute force (oracle/dense.py) before it is written:

  traj_n8_chi16.npz   N=8, depth 8, p=0.3 Haar brickwork trajectory with
                      Const(16) (untruncated at N=8); measurement sites below
                      n-1 so the dense replay applies; the sampled outcome
                      record, probabilities, entropy per layer and the final
                      Schmidt spectra.
  traj_n12_chi4.npz   N=12, depth 12, p=0.1 with Const(4): truncating; the
                      oracle's record, probabilities, entropies, spectra
                      (invariants checked: sum lambda^2 = 1, dims <= 4).
  gates_n6.npz        30 random one- and two-qubit gates on N=6, None
                      truncation: the dense final state and spectra.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
for p in (REPO, REPO / "tensor-net_amd"):
    sys.path.insert(0, str(p))

from oracle import dense  # noqa: E402
from oracle.mps_oracle import VidalMPS, run_trajectory  # noqa: E402
from tnmps.workload import haar4, haar_circuit  # noqa: E402


def pack_layers(layers, n):
    depth = len(layers)
    nu = np.array([len(u) for u, _ in layers], dtype=np.int64)
    uks = np.array([k for u, _ in layers for k, _ in u], dtype=np.int64)
    mats = np.array([m for u, _ in layers for _, m in u], dtype=np.complex128)
    return depth, nu, uks, mats


def pack_record(record):
    nm = np.array([len(r) for r in record], dtype=np.int64)
    ks = np.array([k for r in record for k, _ in r], dtype=np.int64)
    outs = np.array([o for r in record for _, o in r], dtype=np.int64)
    return nm, ks, outs


def pack_spectra(L, chi):
    out = np.zeros((len(L), chi))
    dims = np.array([len(l) for l in L], dtype=np.int64)
    for b, l in enumerate(L):
        out[b, : len(l)] = l
    return out, dims


def traj_fixture(name, n, depth, p, chi, seed, last_site, dense_check):
    c = haar_circuit(n, depth, p, seed)
    if not last_site:
        c.layers = [(u, [k for k in m if k < n - 1]) for u, m in c.layers]
    res = run_trajectory(n, c.layers, ("const", chi), rs=c.draws)
    st = res["state"]
    if dense_check:
        pr, ent, psi = dense.run_dense(n, c.layers, res["record"], n // 2 - 1)
        assert np.max(np.abs(np.array(ent) - np.array(res["entropy"]))) < 1e-12
        pm = [x for layer in res["probs"] for _, x in layer]
        pd = [x for layer in pr for _, x in layer]
        assert np.max(np.abs(np.array(pm) - np.array(pd))) < 1e-12
    else:
        assert max(st.bond_dims()) <= chi
        for l in st.L:
            assert abs(np.sum(l ** 2) - 1) < 1e-13
    depth_, nu, uks, mats = pack_layers(c.layers, n)
    nm, mks, outs = pack_record(res["record"])
    probs = np.array([x for layer in res["probs"] for _, x in layer])
    spec, dims = pack_spectra(st.L, chi)
    np.savez(HERE / name, n=n, depth=depth, p=p, chi=chi, seed=seed, nu=nu, uks=uks,
             mats=mats, nm=nm, mks=mks, outs=outs, probs=probs,
             entropy=np.array(res["entropy"]), spectra=spec, dims=dims)


def gates_fixture():
    n = 6
    rng = np.random.default_rng(606)
    st = VidalMPS(n, None)
    psi = dense.product_zero(n)
    ks, mats, sizes = [], [], []
    for _ in range(30):
        if rng.random() < 0.2:
            k = int(rng.integers(n))
            q, _ = np.linalg.qr(haar4(rng)[:2, :2])
            st.apply_unitary1(k, q)
            psi = dense.apply1(psi, n, k, q)
            m = np.zeros((4, 4), dtype=np.complex128)
            m[:2, :2] = q
            sizes.append(2)
        else:
            k = int(rng.integers(n - 1))
            m = haar4(rng)
            st.apply_unitary2(k, m)
            psi = dense.apply2(psi, n, k, m)
            sizes.append(4)
        ks.append(k)
        mats.append(m)
    spec = [dense.schmidt_values(psi, n, b) for b in range(n - 1)]
    for b in range(n - 1):
        sd = spec[b][spec[b] > 1e-14]
        assert np.max(np.abs(sd - st.L[b][: len(sd)])) < 1e-12
    spec_arr, dims = pack_spectra([s[s > 1e-14] for s in spec], 8)
    np.savez(HERE / "gates_n6.npz", n=n, ks=np.array(ks), sizes=np.array(sizes),
             mats=np.array(mats), psi=psi, spectra=spec_arr, dims=dims)


if __name__ == "__main__":
    traj_fixture("traj_n8_chi16.npz", 8, 8, 0.3, 16, 2024, last_site=False, dense_check=True)
    traj_fixture("traj_n12_chi4.npz", 12, 12, 0.1, 4, 77, last_site=True, dense_check=False)
    gates_fixture()
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)
