"""Admission and continuous batching of the serving loop, on a simulated
executor and a virtual clock (CPU only)."""
import pytest

from diff_envshade_amd.serving import PoissonWorkload, Request, Scheduler, request_bytes, summarize


class VirtualClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def tick(self, dt=1e-4):
        self.t += dt


class SimExecutor:
    """Completes a batch after a duration proportional to its pairs."""

    def __init__(self, clock, slots=3, pairs_per_s=2e10, budget=None):
        self.clock = clock
        self.slots = slots
        self.rate = pairs_per_s
        self.inflight = {}
        self.max_inflight = 0
        self.max_bytes = 0
        self.budget = budget
        self.launches = []
        self._next = 0

    def launch(self, kind, requests):
        assert len(self.inflight) < self.slots
        assert all(r.backward == (kind == "fwdbwd") for r in requests)
        pairs = sum(r.points * r.texels * (3 if r.backward else 1) for r in requests)
        h = self._next
        self._next += 1
        self.inflight[h] = (self.clock() + pairs / self.rate, requests)
        self.max_inflight = max(self.max_inflight, len(self.inflight))
        live = sum(r.bytes for _, rs in self.inflight.values() for r in rs)
        self.max_bytes = max(self.max_bytes, live)
        self.launches.append((kind, [r.rid for r in requests]))
        return h

    def poll(self, h):
        done_t, _ = self.inflight[h]
        if self.clock() >= done_t:
            del self.inflight[h]
            return True
        return False


def _run(reqs, budget, slots=3, max_bypass_s=0.05):
    clk = VirtualClock()
    ex = SimExecutor(clk, slots=slots)
    sch = Scheduler(ex, budget, max_bypass_s=max_bypass_s, clock=clk)
    sch.run(reqs, on_idle=clk.tick)
    return sch, ex


def test_poisson_workload_all_complete_within_budget():
    wl = PoissonWorkload(n_requests=300, rate=3000.0, seed=1)
    reqs = wl.requests()
    budget = 2 << 30
    sch, ex = _run(reqs, budget)
    assert all(r.done_s is not None and r.done_s >= r.arrival_s for r in reqs)
    assert all(r.admitted_s >= r.arrival_s for r in reqs)
    assert ex.max_bytes <= budget
    assert ex.max_inflight <= 3
    s = summarize(reqs, lambda r: r.points)
    assert s["p99_ms"] >= s["p50_ms"] > 0
    assert sch.in_use == 0
    # sizes are log-uniform in [2k, 300k], 30% forward-only
    assert min(r.points for r in reqs) >= 2000 and max(r.points for r in reqs) <= 300000
    frac = sum(not r.backward for r in reqs) / len(reqs)
    assert 0.2 < frac < 0.4


def test_requests_batch_while_others_run():
    reqs = [Request(i, 0.0001 * i, 50000, 512, True) for i in range(40)]
    for r in reqs:
        r.bytes = request_bytes(r.points, r.texels, r.backward)
    sch, ex = _run(reqs, budget=1 << 34)
    # later arrivals are admitted together while earlier batches are in flight
    assert max(len(ids) for _, ids in ex.launches) > 1
    assert ex.max_inflight > 1


def test_fifo_within_kind_and_kinds_never_mixed():
    reqs = []
    for i in range(30):
        r = Request(i, 0.0, 10000 + i, 512, backward=(i % 3 != 0))
        r.bytes = request_bytes(r.points, r.texels, r.backward)
        reqs.append(r)
    sch, ex = _run(reqs, budget=1 << 34, slots=2)
    seen = {True: [], False: []}
    for kind, ids in ex.launches:
        for i in ids:
            seen[kind == "fwdbwd"].append(i)
    for k in seen:
        assert seen[k] == sorted(seen[k])


def test_budget_blocks_and_large_request_is_not_starved():
    budget = 100 * (1 << 20)
    big = Request(0, 0.0, 300000, 2048, True)
    big.bytes = 90 * (1 << 20)
    small = []
    for i in range(1, 200):
        r = Request(i, 0.0002 * i, 5000, 512, True)
        r.bytes = 20 * (1 << 20)
        small.append(r)
    # the big one arrives while small ones hold memory
    big.arrival_s = 0.001
    reqs = small + [big]
    sch, ex = _run(reqs, budget, slots=4, max_bypass_s=0.002)
    assert ex.max_bytes <= budget
    order = [i for _, ids in ex.launches for i in ids]
    # it is admitted long before the stream of small requests ends
    assert order.index(0) < len(order) - 50
    assert big.done_s is not None


def test_request_above_budget_is_rejected():
    r = Request(0, 0.0, 300000, 2048, True)
    r.bytes = request_bytes(r.points, r.texels, r.backward)
    with pytest.raises(RuntimeError, match="above the whole HBM budget"):
        _run([r], budget=1 << 20)


def test_burst_arrivals_come_together_and_stay_ordered():
    wl = PoissonWorkload(n_requests=200, rate=1000.0, seed=3, burst=(50, 40))
    a = wl.arrivals
    assert (a[50:90] == a[50]).all()
    assert (a[1:] >= a[:-1]).all()
    assert a[90] > a[89]


def test_max_overlap():
    from diff_envshade_amd.serving import max_overlap
    assert max_overlap([]) == 0
    assert max_overlap([(0, 1), (1, 2)]) == 1  # touching intervals do not overlap
    assert max_overlap([(0, 3), (1, 2), (1.5, 4)]) == 3


class WsSimExecutor(SimExecutor):
    """A simulated executor with preallocated batch workspaces."""

    def __init__(self, clock, slots, ws_bytes):
        super().__init__(clock, slots=slots)
        self.fixed_bytes = slots * ws_bytes
        self.batch_workspace_bytes = ws_bytes
        self.batch_ws = []

    def launch(self, kind, requests):
        self.batch_ws.append(sum(r.ws_bytes for r in requests))
        return super().launch(kind, requests)


def test_preallocated_workspaces_are_charged_up_front_and_bound_each_batch():
    wl = PoissonWorkload(n_requests=300, rate=1e6, seed=4)
    reqs = wl.requests()
    assert all(0 < r.ws_bytes < r.bytes for r in reqs)
    clk = VirtualClock()
    ws = 200 << 20
    ex = WsSimExecutor(clk, slots=3, ws_bytes=ws)
    budget = 3 * ws + (300 << 20)
    sch = Scheduler(ex, budget, clock=clk)
    assert sch.budget == 300 << 20
    sch.run(reqs, on_idle=clk.tick)
    assert all(r.done_s is not None for r in reqs)
    assert max(ex.batch_ws) <= ws
    assert sch.peak_in_use <= 300 << 20
    # a request is charged its outputs and gradients, not its workspace share
    assert sch.charge(reqs[0]) == reqs[0].bytes - reqs[0].ws_bytes
    with pytest.raises(RuntimeError, match="more than the HBM budget"):
        Scheduler(WsSimExecutor(clk, slots=3, ws_bytes=ws), 3 * ws, clock=clk)
