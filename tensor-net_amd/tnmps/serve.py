"""Continuous-batching trajectory server over one engine.

Each request is one trajectory: an ``MPS`` of ``n`` qubits with
``Const(chi)`` truncation driven through ``depth`` bilayers
(``circuit::apply_bilayer``, circuit.rs:624-659) of its own circuit, with
``Meas::Rand`` outcomes drawn from its own uniform stream, as the
reference's trajectory drivers do one MPS per thread
(workspace/src/haar_coev_probs.rs:33-83, 162-203).

A request either carries its circuit (``layers`` and ``draws``) or is
generated: ``gen_seed`` and ``traj_id`` fix its Haar gates, measurement
sites and draws, which the library makes on the device
(``tn_batch_bilayer_gen``), so no gate crosses the host-device boundary.

The server keeps every admitted trajectory resident on the device and, per
tick, advances all of them by one layer in lock step (one
``batch_bilayer`` call for carried circuits, one ``batch_bilayer_gen``
call per generator seed), so their gates and sweep steps share launches.
Requests that have arrived are admitted in arrival order whenever the HBM
held stays within the budget: each worker engine's update workspace (its
cap, set from the budget when the engines are made) plus the resident
trajectories' Gamma and lambda; finished trajectories leave at once and free
their memory.

With ``workers=W > 1`` the server runs W such lock-step loops at once, each
on its own engine (stream) and host thread.  A lock-step tick lasts as long
as its slowest SVD, so independent loops keep one worker's slow tail from
holding back the others' trajectories.  The HBM budget is shared.  Requests
go to workers by one of two rules (``dispatch``):
- ``"least_loaded"`` (default): one arrival-ordered queue; a worker admits
  the next arrived request at its tick boundary only while no other worker
  holds less outstanding work (an idle worker holds none).  A request's work
  is estimated as its remaining layers x n x chi_eff^3, chi_eff the largest
  bond it can reach: the cost of a layer's SVDs.
- ``"round_robin"``: request i of the arrival order goes to worker i mod W.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from .mps import Engine, MPS


def layer_cost(n: int, chi: int) -> float:
    """Relative cost of one layer of an n-qubit Const(chi) trajectory: about
    2n two-site SVDs of order chi_eff^3, chi_eff the largest reachable bond."""
    ce = min(chi, 2 ** min(n // 2, 40))
    return float(n) * float(ce) ** 3


GIB = float(1 << 30)


def workspace_bytes(chi_max: int, ws_cap: float) -> float:
    """Device bytes an engine's update workspace can reach under a
    TNMPS_WS_BYTES cap of ws_cap: its slots (X and W of 2chi x 2chi complex
    plus per-column arrays, as the library counts them; at least one) and
    the row-split partial Gram matrices of up to 256 slots (one per CU)."""
    ld = 2 * chi_max
    per = 16.0 * (2.0 * ld * ld) + 64.0 * ld + 4096.0
    slots = max(1.0, float(int(ws_cap // per)))
    gpart = max(((ld + bw - 1) // bw // 2 + 1) * 8 * 4 * bw * bw for bw in (8, 16)) * 16.0
    return slots * per + min(slots, 256.0) * gpart


def mps_bytes(n: int, chi: int) -> int:
    """Device bytes of an n-qubit MPS with bonds capped at chi (the library's
    allocation: Gamma slots sized for the largest reachable bond dimension)."""
    b = [min(chi, 2 ** min(k + 1, n - k - 1, 40)) for k in range(n - 1)]
    g = 0
    for k in range(n):
        left = 1 if k == 0 else b[k - 1]
        right = 1 if k == n - 1 else b[k]
        g += left * 2 * right
    return 16 * g + 8 * sum(b) + 4 * (n + 1)


@dataclass
class Request:
    """One trajectory request: its circuit and draws, or its generator seed,
    trajectory id and depth; and its arrival time (s)."""
    rid: int
    n: int
    p: float
    chi: int
    layers: Optional[list] = None   # [(unis, meas_sites)] of length depth
    draws: Sequence[float] = ()
    arrival: float = 0.0
    prec: str = "f64"
    gen_seed: Optional[int] = None  # generated on the device when set
    traj_id: int = 0
    gen_depth: int = 0
    # filled by the server
    admitted: Optional[float] = None
    done: Optional[float] = None
    record: list = field(default_factory=list)
    probs: list = field(default_factory=list)
    entropy: list = field(default_factory=list)

    @property
    def depth(self) -> int:
        return self.gen_depth if self.gen_seed is not None else len(self.layers)

    @property
    def latency(self) -> Optional[float]:
        return None if self.done is None else self.done - self.arrival


class Server:
    """Admit, advance in lock step, retire.  ``hbm_budget`` bounds the bytes of
    the resident trajectories' state; ``max_active`` bounds their number (per
    worker); ``workers`` lock-step loops run at once, one engine each."""

    def __init__(self, chi_max: int = 256, hbm_budget: float = 64e9, max_active: int = 1024,
                 log: Optional[Callable[[str], None]] = None, workers: int = 1,
                 dispatch: str = "least_loaded", ws_cap: Optional[float] = None):
        """``ws_cap``: each worker engine's workspace cap (TNMPS_WS_BYTES),
        by default a quarter of the budget over the workers, at most 24 GiB;
        what the workspaces can reach counts against the budget from the
        start."""
        if dispatch not in ("least_loaded", "round_robin"):
            raise ValueError(f"dispatch {dispatch!r}")
        self.dispatch = dispatch
        nw = max(1, workers)
        if ws_cap is None:
            ws_cap = min(24 * GIB, 0.25 * hbm_budget / nw)
        import os
        old = os.environ.get("TNMPS_WS_BYTES")
        os.environ["TNMPS_WS_BYTES"] = str(int(ws_cap))
        try:
            self.engines = [Engine(chi_max) for _ in range(nw)]
        finally:
            if old is None:
                del os.environ["TNMPS_WS_BYTES"]
            else:
                os.environ["TNMPS_WS_BYTES"] = old
        self.ws_held = nw * workspace_bytes(chi_max, ws_cap)
        self._load = [0.0] * len(self.engines)  # outstanding work per worker
        self.engine = self.engines[0]
        self.chi_max = chi_max
        self.hbm_budget = hbm_budget
        self.max_active = max_active
        self.ticks = 0
        self.layers_done = 0
        self.peak_held = 0.0   # largest bytes held while more than one was active
        self.log = log
        self._lock = threading.Lock()
        self._held = self.ws_held
        self._nactive = 0

    def stats(self):
        """(updates, sweeps) summed over the server's engines."""
        u = s = 0
        for e in self.engines:
            a, b = e.stats()
            u += a
            s += b
        return u, s

    # ------------------------------------------------------------ admission
    def _try_admit(self, r: Request, eng: Engine, active: list, t0: float, clock) -> bool:
        """Admit r onto eng if the budget allows (always when nothing is
        resident); under the lock."""
        need = mps_bytes(r.n, r.chi)
        if self._nactive and self._held + need > self.hbm_budget:
            return False
        st = eng.new_qubits(r.n, ("const", r.chi))
        if r.prec != "f64":
            st.set_precision(r.prec)
        r.admitted = clock() - t0
        r.entropy = [0.0]  # product state: entropy before layer 1
        active.append([r, st, iter(r.draws), 0])
        self._held += need
        self._nactive += 1
        if self._nactive > 1:
            self.peak_held = max(self.peak_held, self._held)
        return True

    # ----------------------------------------------------------------- tick
    def _tick(self, eng: Engine, active: list, t0: float, clock) -> list:
        """Advance every trajectory of ``active`` by one layer on ``eng``;
        retire finished ones and return the rest."""
        res = [None] * len(active)
        carried = [i for i, a in enumerate(active) if a[0].gen_seed is None]
        if carried:
            got = eng.batch_bilayer(
                [active[i][1] for i in carried],
                [active[i][0].layers[active[i][3]] for i in carried],
                draws=[active[i][2] for i in carried],
                ent_bonds=[active[i][0].n // 2 - 1 for i in carried])
            for i, g in zip(carried, got):
                res[i] = g
        for seed in sorted({a[0].gen_seed for a in active if a[0].gen_seed is not None}):
            idx = [i for i, a in enumerate(active) if a[0].gen_seed == seed]
            got = eng.batch_bilayer_gen(
                [active[i][1] for i in idx], seed,
                [active[i][0].traj_id for i in idx], [active[i][3] for i in idx],
                [active[i][0].p for i in idx],
                ent_bonds=[active[i][0].n // 2 - 1 for i in idx])
            for i, g in zip(idx, got):
                res[i] = g
        t_layer = clock() - t0
        keep = []
        done = []
        for a, (rec, prs, ent) in zip(active, res):
            r = a[0]
            r.record.append(rec)
            r.probs.append(prs)
            r.entropy.append(ent)
            a[3] += 1
            if a[3] >= r.depth:
                r.done = t_layer
                a[1].free()
                done.append(r)
            else:
                keep.append(a)
        with self._lock:
            self.ticks += 1
            self.layers_done += len(active)
            for r in done:
                self._held -= mps_bytes(r.n, r.chi)
                self._nactive -= 1
            if self.log and self.ticks % 10 == 0:
                self.log(f"tick {self.ticks}: {self._nactive} active")
        if self.log:
            for r in done:
                self.log(f"t={r.done:.1f}s request {r.rid} (n={r.n}, chi={r.chi}) done, "
                         f"latency {r.latency:.2f}s")
        return keep

    # ------------------------------------------------------------------ run
    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            realtime: bool = True) -> List[Request]:
        """Serve ``requests`` (arrival times relative to the start).  With
        ``realtime`` the server waits for arrivals on the wall clock;
        otherwise every request counts as arrived at its time stamp as soon
        as the server is idle or at a tick boundary past it (one worker
        only)."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = clock()
        if len(self.engines) == 1:
            self._loop(self.engine, pending, t0, clock, realtime)
            return requests
        if not realtime:
            raise ValueError("more than one worker needs realtime=True")
        W = len(self.engines)
        errors: list = []
        shared = {"queue": pending, "next": 0}

        def work(w):
            try:
                if self.dispatch == "round_robin":
                    self._loop(self.engines[w], pending[w::W], t0, clock, True)
                else:
                    self._loop_shared(w, shared, t0, clock)
            except Exception as exc:  # surfaced after the join
                errors.append(exc)

        th = [threading.Thread(target=work, args=(w,)) for w in range(W)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        if errors:
            raise errors[0]
        return requests

    def _loop(self, eng: Engine, pending: list, t0: float, clock, realtime: bool) -> None:
        """One lock-step loop over ``pending`` (in arrival order) on ``eng``."""
        active: list = []           # [(req, state, draws iterator, next layer)]
        qi = 0
        while qi < len(pending) or active:
            now = clock() - t0
            if not active and qi < len(pending) and pending[qi].arrival > now:
                if realtime:
                    time.sleep(min(pending[qi].arrival - now, 0.05))
                    continue
                now = pending[qi].arrival
            # admit in arrival order while memory and slots allow
            with self._lock:
                while (qi < len(pending) and pending[qi].arrival <= now + 1e-12
                       and len(active) < self.max_active):
                    if not self._try_admit(pending[qi], eng, active, t0, clock):
                        break
                    qi += 1
            if not active:
                if realtime and qi < len(pending):
                    time.sleep(0.005)   # budget held by other workers
                continue
            active = self._tick(eng, active, t0, clock)


    def _loop_shared(self, w: int, shared: dict, t0: float, clock) -> None:
        """Worker w's lock-step loop over the shared arrival-ordered queue,
        least-loaded rule (see the module docstring)."""
        eng = self.engines[w]
        queue = shared["queue"]
        active: list = []
        while True:
            now = clock() - t0
            with self._lock:
                while (shared["next"] < len(queue)
                       and queue[shared["next"]].arrival <= now + 1e-12
                       and len(active) < self.max_active
                       and may_admit(self._load, w)):
                    r = queue[shared["next"]]
                    if not self._try_admit(r, eng, active, t0, clock):
                        break
                    shared["next"] += 1
                    self._load[w] += r.depth * layer_cost(r.n, r.chi)
                drained = shared["next"] >= len(queue)
            if not active:
                if drained:
                    return
                nxt = queue[min(shared["next"], len(queue) - 1)].arrival
                time.sleep(min(max(nxt - now, 0.0), 0.005) or 0.001)
                continue
            active = self._tick(eng, active, t0, clock)
            with self._lock:
                self._load[w] = sum((a[0].depth - a[3]) * layer_cost(a[0].n, a[0].chi)
                                    for a in active)


def may_admit(load: Sequence[float], w: int) -> bool:
    """Least-loaded rule: worker w may take the next request when no other
    worker holds less outstanding work."""
    return load[w] <= min(load)


def percentile(xs: Sequence[float], q: float) -> float:
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else float("nan")
