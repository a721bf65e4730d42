"""Serving loop: a stream of solve requests into the batch engine.

Requests arrive at their own times (a seeded Poisson process in the
benchmark).  Arrived requests queue first-in first-out; the head of the
queue is admitted as soon as the engine has a free slot and enough free HBM
budget for its footprint (``Batch.admit`` refuses otherwise), later requests
wait behind it.  Every loop iteration advances all resident cases together
(continuous batching: a case joins at the next step after its admission and
leaves the step after its last), then retires finished cases.  An iteration
runs as many steps back to back as fit before the next arrival, at the
measured step time, and stops early when a case finishes, so arrivals are
admitted within about one step and the device is not idled by per-step
host round trips.
Latency of a request = completion time - arrival time.  The reference has
no serving path (one case per ``mpirun``, ouxsbli/case.py:180-201).
"""
from __future__ import annotations

import collections
import concurrent.futures
import dataclasses
import os
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from ._native import OxError
from .batch import Batch, Busy
from .cases import CaseSpec, dhit_spec, initial_state, khi_spec


@dataclasses.dataclass
class Request:
    rid: int
    spec: CaseSpec
    nsteps: int
    arrival: float                       # seconds after the stream starts
    Q0: Optional[np.ndarray] = None      # initial physical state (the payload); None: built on the device
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[np.ndarray] = None
    error: Optional[str] = None          # set when the request was rejected (it can never be admitted)

    @property
    def cell_stages(self) -> float:
        return float(self.spec.interior_cells) * self.spec.stages * self.nsteps


def khi_stream(n_requests: int, rate_hz: float = 20.0, seed: int = 20261207, n_min: int = 64, n_max: int = 256,
               multiple: int = 16, nsteps: int = 200, precision: str = "fp32_bulk",
               with_states: bool = True) -> List[Request]:
    """The Kelvin-Helmholtz request stream: exponential inter-arrival times
    (mean 1/rate_hz), interior size n uniform over the multiples of
    ``multiple`` in [n_min, n_max], scheme KEEP or SLAU and ORDER 4 or 6 at
    random, ``nsteps`` RK3 steps each, seeded."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for r in range(n_requests):
        t += float(rng.exponential(1.0 / rate_hz))
        n = int(rng.integers(n_min // multiple, n_max // multiple + 1)) * multiple
        scheme = str(rng.choice(["KEEP", "SLAU"]))
        order = int(rng.choice([4, 6]))
        spec = khi_spec(n, order, scheme, precision, seed=seed + 1 + r)
        reqs.append(Request(r, spec, nsteps, t))
    if with_states:
        # each request's state comes from its own seeded generator, so the
        # threads do not change what is generated
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            for r, Q in zip(reqs, ex.map(lambda q: initial_state(q.spec), reqs)):
                r.Q0 = Q
    return reqs


def dhit_stream(n_requests: int, rate_hz: float = 4.0, seed: int = 20261208, n_min: int = 64, n_max: int = 192,
                multiple: int = 32, nsteps: int = 50, precision: str = "fp32_bulk",
                with_states: bool = False) -> List[Request]:
    """An LES request stream: decaying isotropic turbulence (the LES
    scheme: Hybrid KEEP/SLAU ORDER 6, SMS model, RK4) on interiors uniform
    over the multiples of ``multiple`` in [n_min, n_max], each with its own
    spectral seed, exponential inter-arrival times (mean 1/rate_hz).  By
    default the requests carry parameters only and the engine synthesises
    each initial field on the device."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for r in range(n_requests):
        t += float(rng.exponential(1.0 / rate_hz))
        n = int(rng.integers(n_min // multiple, n_max // multiple + 1)) * multiple
        spec = dhit_spec(n, precision, seed=seed + 1 + r)
        reqs.append(Request(r, spec, nsteps, t, Q0=initial_state(spec) if with_states else None))
    return reqs


@dataclasses.dataclass
class ServeReport:
    requests: List[Request]
    wall_s: float
    batch_stats: Dict
    steps: int
    host_s: Dict = dataclasses.field(default_factory=dict)  # wall seconds in admit / run / retire calls

    def served(self) -> List[Request]:
        return [r for r in self.requests if r.error is None]

    def latencies(self) -> np.ndarray:
        return np.array([r.finished - r.arrival for r in self.served()])

    def summary(self) -> Dict:
        lat = self.latencies()
        if lat.size == 0:
            lat = np.array([float("nan")])
        cs = sum(r.cell_stages for r in self.served())
        return {
            "requests": len(self.requests),
            "rejected": sum(1 for r in self.requests if r.error is not None),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "max_latency_s": float(lat.max()),
            "wall_s": self.wall_s,
            "Mcell_updates_per_s_stage": cs / self.wall_s / 1e6,
            "device_busy_Mcell_updates_per_s_stage": cs / (self.batch_stats["device_ms"] * 1e-3) / 1e6
            if self.batch_stats["device_ms"] > 0 else None,
            "batch_steps": self.steps,
            "graph_captures": self.batch_stats["graph_captures"],
            "graph_replays": self.batch_stats["graph_replays"],
            **{f"{k}_s": round(v, 4) for k, v in self.host_s.items()},
        }


def serve(requests: List[Request], hbm_budget: int, max_cases: int = 64, use_graph: bool = True,
          autotune: bool = True, pretune: bool = False, keep_results: bool = True,
          clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
          batch: Optional[Batch] = None, max_run: int = 1024) -> ServeReport:
    """Runs the stream to completion and returns the per-request timings.

    With ``autotune`` the engine tunes a shape's segment lengths when it first
    admits it, inside that request's latency, and keeps the result for later
    requests of the same shape (an engine passed in as ``batch`` keeps what
    earlier streams tuned).  ``pretune`` instead fills the cache for every
    shape of the stream before the clock starts, which a live server cannot
    do; it is off by default.  The report's engine statistics cover this
    stream only."""
    own = batch is None
    if own:
        batch = Batch(hbm_budget, max_cases=max_cases, use_graph=use_graph, autotune=autotune)
    try:
        # a request without a payload is built on the device at admission
        # when its case has a device initial condition, on the host otherwise
        for r in requests:
            if r.Q0 is None and r.spec.device_ic() is None:
                r.Q0 = initial_state(r.spec)
        if pretune and autotune:
            seen = set()
            for r in requests:
                s = r.spec
                key = (s.precision, s.scheme, s.order, s.nx, s.ny, s.nz)
                if key not in seen:
                    seen.add(key)
                    batch.tune(s, r.Q0 if r.Q0 is not None else initial_state(s))
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting: collections.deque = collections.deque()
        by_id: Dict[int, Request] = {}
        active = 0
        steps = 0
        step_s = None  # running estimate of the wall time of one batch step
        host = {"admit": 0.0, "run": 0.0, "retire": 0.0}
        stats0 = batch.stats()
        t0 = clock()
        while pending or waiting or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            while waiting:
                r = waiting[0]
                ta = clock()
                try:
                    cid = batch.admit(r.spec, r.Q0, r.nsteps)
                except Busy as e:
                    host["admit"] += clock() - ta
                    if active == 0:
                        # nothing resident can free a slot or budget: the
                        # request does not fit an empty engine
                        r.error = f"not admissible into an empty engine: {e}"
                        waiting.popleft()
                        continue
                    break
                except OxError as e:
                    # permanently refused (footprint over the whole budget,
                    # unsupported configuration): reject, serve the rest
                    host["admit"] += clock() - ta
                    r.error = str(e)
                    waiting.popleft()
                    continue
                host["admit"] += clock() - ta
                r.admitted = clock() - t0
                by_id[cid] = r
                waiting.popleft()
                active += 1
            if active == 0:
                if pending:
                    sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            if pending and step_s is not None:
                k = int((pending[0].arrival - (clock() - t0)) / step_s)
                k = max(1, min(max_run, k))
            elif pending:
                k = 1
            else:
                k = max_run  # stops at the first finish, which is when a waiting request can enter
            t_run = clock()
            ran, active = batch.run(k)
            host["run"] += clock() - t_run
            per = (clock() - t_run) / max(ran, 1)
            step_s = per if step_s is None else 0.7 * step_s + 0.3 * per
            steps += ran
            for cid in batch.done():
                r = by_id.pop(cid)
                tr = clock()
                Q = batch.retire(cid)
                host["retire"] += clock() - tr
                r.finished = clock() - t0
                if keep_results:
                    r.result = Q
        wall = clock() - t0
        stats = dict(batch.stats())
        for k in ("device_ms", "graph_captures", "graph_replays", "steps", "cell_stages"):
            if k in stats and k in stats0:
                stats[k] = stats[k] - stats0[k]
        return ServeReport(requests, wall, stats, steps, host)
    finally:
        if own:
            batch.close()


def main(argv=None) -> int:
    """Runs a request stream (Kelvin-Helmholtz, or LES isotropic turbulence)
    and prints one JSON line with the latency percentiles, throughput and
    engine statistics."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--stream", default="khi", choices=["khi", "dhit"])
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=None, help="mean arrivals per second (khi 20, dhit 4)")
    ap.add_argument("--n-min", type=int, default=64)
    ap.add_argument("--n-max", type=int, default=None, help="largest interior (khi 256, dhit 192)")
    ap.add_argument("--steps", type=int, default=None, help="steps per request (khi 200, dhit 50)")
    ap.add_argument("--precision", default="fp32_bulk", choices=["fp64", "fp32_bulk"])
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--slots", type=int, default=16)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--seed", type=int, default=None)
    a = ap.parse_args(argv)
    if a.stream == "khi":
        reqs = khi_stream(a.requests, rate_hz=a.rate or 20.0, seed=a.seed or 20261207, n_min=a.n_min,
                          n_max=a.n_max or 256, nsteps=a.steps or 200, precision=a.precision, with_states=False)
    else:
        reqs = dhit_stream(a.requests, rate_hz=a.rate or 4.0, seed=a.seed or 20261208, n_min=a.n_min,
                           n_max=a.n_max or 192, nsteps=a.steps or 50, precision=a.precision)
    rep = serve(reqs, hbm_budget=int(a.budget_gib * (1 << 30)), max_cases=a.slots, use_graph=not a.no_graph,
                keep_results=False)
    print(json.dumps({**rep.summary(), "engine": rep.batch_stats}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
