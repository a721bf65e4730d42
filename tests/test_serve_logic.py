"""Host logic of the serving loop (ouxsbli_amd/serve.py) against a simulated
engine and clock: FIFO admission by slots and HBM footprint, continuous
batching at step granularity, latency bookkeeping.  No GPU."""
import numpy as np
import pytest

from ouxsbli_amd import serve
from ouxsbli_amd.batch import Busy
from ouxsbli_amd.cases import khi_spec


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 0.0)


class FakeBatch:
    """Admission by slots and a byte budget; one step costs `step_s` of the
    simulated clock; a case finishes after its step count."""

    def __init__(self, clock, budget, slots, step_s=0.01):
        self.clock, self.budget, self.slots, self.step_s = clock, budget, slots, step_s
        self.active, self.finished, self.used, self.next = {}, [], 0, 1
        self.sizes = {}
        self.log = []
        self.tuned = []

    @staticmethod
    def bytes(spec):
        return spec.nx * spec.ny * spec.nz

    def tune(self, spec, Q):
        self.tuned.append((spec.nx, spec.scheme, spec.order))

    def admit(self, spec, Q, nsteps):
        if len(self.active) + len(self.finished) >= self.slots or self.used + self.bytes(spec) > self.budget:
            raise Busy("full")
        cid = self.next
        self.next += 1
        self.active[cid] = [nsteps, Q]
        self.sizes[cid] = self.bytes(spec)
        self.used += self.bytes(spec)
        self.log.append(("admit", cid, self.clock()))
        return cid

    def step(self):
        self.clock.t += self.step_s
        self.log.append(("step", None, self.clock()))
        for cid in list(self.active):
            self.active[cid][0] -= 1
            if self.active[cid][0] == 0:
                self.finished.append(cid)
        for cid in self.finished:
            self.active.pop(cid, None)
        return len(self.active)

    def run(self, k):
        """Up to k steps, stopping after the first one in which a case finishes."""
        n = 0
        while n < k and self.active:
            before = len(self.finished)
            self.step()
            n += 1
            if len(self.finished) > before:
                break
        return n, len(self.active)

    def done(self):
        return list(self.finished)

    def retire(self, cid):
        self.finished.remove(cid)
        self.used -= self.sizes.pop(cid)
        return np.zeros(1)

    def stats(self):
        return {"device_ms": 0.0, "graph_captures": 0, "graph_replays": 0}


def _reqs(arrivals, sizes, nsteps):
    out = []
    for r, (t, n, k) in enumerate(zip(arrivals, sizes, nsteps)):
        out.append(serve.Request(r, khi_spec(n, 4, "KEEP"), k, t, Q0=np.zeros(1)))
    return out


def test_stream_is_seeded_and_in_range():
    a = serve.khi_stream(16, with_states=False, seed=7)
    b = serve.khi_stream(16, with_states=False, seed=7)
    assert [(r.arrival, r.spec.nx, r.spec.scheme, r.spec.order) for r in a] == \
           [(r.arrival, r.spec.nx, r.spec.scheme, r.spec.order) for r in b]
    for r in a:
        n = r.spec.nx - r.spec.order
        assert 64 <= n <= 256 and n % 16 == 0
        assert r.spec.scheme in ("KEEP", "SLAU") and r.spec.order in (4, 6)
        assert r.spec.rk == 3 and r.spec.visc == "Euler"
    gaps = np.diff([0.0] + [r.arrival for r in serve.khi_stream(400, rate_hz=20.0, with_states=False)])
    assert abs(gaps.mean() - 0.05) < 0.01  # mean inter-arrival 1/rate


def test_les_stream_is_seeded_parameters_only():
    a = serve.dhit_stream(12, seed=5)
    b = serve.dhit_stream(12, seed=5)
    assert [(r.arrival, r.spec.nx, r.spec.seed) for r in a] == [(r.arrival, r.spec.nx, r.spec.seed) for r in b]
    assert len({r.spec.seed for r in a}) == len(a)  # each request its own field
    for r in a:
        n = r.spec.nx - r.spec.order
        assert 64 <= n <= 192 and n % 32 == 0
        assert r.Q0 is None and r.spec.device_ic() is not None
        assert (r.spec.visc, r.spec.scheme, r.spec.order, r.spec.rk) == ("LES", "Hybrid", 6, 4)


def test_all_requests_complete_with_fifo_admission():
    clock = FakeClock()
    fb = FakeBatch(clock, budget=10 ** 9, slots=2)
    reqs = _reqs([0.0, 0.0, 0.001, 0.5], [8, 8, 8, 8], [3, 5, 2, 1])
    rep = serve.serve(reqs, hbm_budget=0, batch=fb, clock=clock, sleep=clock.sleep, pretune=False)
    assert all(r.finished is not None for r in rep.requests)
    # two slots: request 2 waits for a slot, request 3 arrives later
    admits = [e for e in fb.log if e[0] == "admit"]
    assert [e[1] for e in admits] == [1, 2, 3, 4]
    r0, r1, r2, r3 = rep.requests
    assert r0.finished == pytest.approx(0.03) and r1.finished == pytest.approx(0.05)
    assert r2.admitted >= r0.finished - 1e-12      # admitted when the first slot freed
    assert r2.finished == pytest.approx(r2.admitted + 0.02)
    assert r3.admitted >= 0.5
    lat = rep.latencies()
    assert np.all(lat > 0)
    s = rep.summary()
    assert s["requests"] == 4 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_budget_blocks_head_of_line():
    clock = FakeClock()
    small, big = khi_spec(8, 4, "KEEP"), khi_spec(24, 4, "KEEP")
    budget = FakeBatch.bytes(big) + FakeBatch.bytes(small) - 1  # big and small never together
    fb = FakeBatch(clock, budget=budget, slots=8)
    reqs = [serve.Request(0, big, 4, 0.0, Q0=np.zeros(1)), serve.Request(1, small, 1, 0.0, Q0=np.zeros(1))]
    # only the head fits; the loop must not reorder the queue past it
    rep = serve.serve(reqs, hbm_budget=0, batch=fb, clock=clock, sleep=clock.sleep, pretune=False)
    r0, r1 = rep.requests
    assert r1.admitted >= r0.finished - 1e-12


def test_pretune_visits_each_shape_once():
    clock = FakeClock()
    fb = FakeBatch(clock, budget=10 ** 9, slots=4)
    reqs = _reqs([0.0, 0.1, 0.2], [8, 8, 16], [1, 1, 1])
    serve.serve(reqs, hbm_budget=0, batch=fb, clock=clock, sleep=clock.sleep, autotune=True, pretune=True)
    assert len(fb.tuned) == 2


def test_runs_back_to_back_between_arrivals():
    """With nothing arriving for a while the loop runs many steps per
    iteration (one host round trip), and a request arriving mid-run is still
    admitted within a couple of steps of its arrival."""
    clock = FakeClock()
    fb = FakeBatch(clock, budget=10 ** 9, slots=4, step_s=0.01)
    reqs = _reqs([0.0, 0.505], [8, 8], [100, 3])
    rep = serve.serve(reqs, hbm_budget=0, batch=fb, clock=clock, sleep=clock.sleep, pretune=False)
    r0, r1 = rep.requests
    assert r0.finished == pytest.approx(1.0)
    assert 0.505 <= r1.admitted <= 0.505 + 2 * 0.01 + 1e-9
    assert r1.finished == pytest.approx(r1.admitted + 0.03)
    assert rep.steps == 100


def test_request_that_never_fits_is_rejected():
    """A head-of-line request whose footprint exceeds the whole budget is
    refused even by an empty engine; the loop rejects it and serves the
    rest instead of spinning (engines that report it as Busy and as a
    permanent error are both handled)."""
    from ouxsbli_amd._native import OxError

    class PermanentBatch(FakeBatch):
        def admit(self, spec, Q, nsteps):
            if self.bytes(spec) > self.budget:
                raise OxError(-1, "footprint exceeds the whole HBM budget")
            return super().admit(spec, Q, nsteps)

    for cls in (FakeBatch, PermanentBatch):
        clock = FakeClock()
        small, big = khi_spec(8, 4, "KEEP"), khi_spec(24, 4, "KEEP")
        fb = cls(clock, budget=FakeBatch.bytes(big) - 1, slots=4)
        reqs = [serve.Request(0, big, 4, 0.0, Q0=np.zeros(1)), serve.Request(1, small, 2, 0.01, Q0=np.zeros(1))]
        rep = serve.serve(reqs, hbm_budget=0, batch=fb, clock=clock, sleep=clock.sleep, pretune=False)
        r0, r1 = rep.requests
        assert r0.error is not None and r0.finished is None
        assert r1.error is None and r1.finished == pytest.approx(r1.admitted + 0.02)
        s = rep.summary()
        assert s["requests"] == 2 and s["rejected"] == 1
        assert clock.t < 1.0
