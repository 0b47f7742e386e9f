"""Host-side logic of the serving loop (uclchem_amd/serving.py) on the CPU: the HBM-budget
slot planner and the seeded Poisson request stream.  The per-slot byte count, which the
library reports for a loaded network, is replaced by fixed sizes here."""
import numpy as np
import pytest

from uclchem_amd import serving
from uclchem_amd.gpu import GpuError
from uclchem_amd.params import INDEX


@pytest.fixture
def fake_sizes(monkeypatch):
    sizes = {"small": 100.0, "default": 300.0, "big": 1000.0}
    monkeypatch.setattr(serving.RequestQueue, "slot_bytes", staticmethod(lambda net, lu_bits=64: sizes[net]))
    return {k: k for k in sizes}          # "networks": the planner only passes them to slot_bytes


def test_plan_slots_shares_cus_within_budget(fake_sizes):
    slots = serving.plan_slots(fake_sizes, n_cu=256, hbm_budget_bytes=1e9)
    assert set(slots) == set(fake_sizes)
    assert sum(slots.values()) <= 256                     # every queue's workgroups resident together
    assert max(slots.values()) - min(slots.values()) <= 1  # equal weights: equal shares
    w = serving.plan_slots(fake_sizes, 256, 1e9, weights={"small": 1.0, "default": 1.0, "big": 2.0})
    assert w["big"] >= 2 * w["small"] - 1


def test_plan_slots_trims_the_largest_consumer_to_fit(fake_sizes):
    budget = 40_000.0
    slots = serving.plan_slots(fake_sizes, n_cu=256, hbm_budget_bytes=budget)
    used = sum(slots[n] * {"small": 100.0, "default": 300.0, "big": 1000.0}[n] for n in slots)
    assert used <= budget and min(slots.values()) >= 1
    tight = serving.plan_slots(fake_sizes, n_cu=256, hbm_budget_bytes=1500.0)   # one slot each fits
    assert min(tight.values()) >= 1
    assert sum(tight[n] * {"small": 100.0, "default": 300.0, "big": 1000.0}[n] for n in tight) <= 1500.0
    with pytest.raises(GpuError):
        serving.plan_slots(fake_sizes, n_cu=256, hbm_budget_bytes=1000.0)


def test_poisson_requests_seeded_and_well_formed():
    names = ["a", "b", "c"]
    r1 = serving.poisson_requests(400, 20.0, names)
    r2 = serving.poisson_requests(400, 20.0, names)
    assert [(t, n) for t, n, _ in r1] == [(t, n) for t, n, _ in r2]
    t = np.array([x[0] for x in r1])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    assert abs(np.mean(np.diff(t)) - 1 / 20.0) < 0.25 / 20.0          # mean gap 1/rate
    assert {x[1] for x in r1} == set(names)
    t_end = np.array([x[2][INDEX["finaltime"]] for x in r1])
    assert t_end.min() >= 1e3 and t_end.max() <= 1e7


class _FakeQueue:
    """Stands in for one device request queue: a request finishes `delay` seconds after it
    is posted, with a result that depends only on its parameters."""

    def __init__(self, net, capacity, n_slots, lu_bits=64, idle_timeout_s=600.0):
        self.posted, self.open, self.delay = [], True, 0.02

    def submit(self, params):
        import time
        self.posted.append((time.perf_counter(), np.array(params, dtype=float)))
        return len(self.posted) - 1

    def done(self, request):
        import time
        return time.perf_counter() - self.posted[request][0] >= self.delay

    def result(self, request):
        p = self.posted[request][1]
        return np.full(3, p[INDEX["finaltime"]]), 0, np.zeros(8, np.int32)

    def close(self):
        self.open = False


def test_stream_replay_returns_every_request_with_its_own_result(monkeypatch):
    monkeypatch.setattr(serving, "RequestQueue", _FakeQueue)
    names = ["small", "default", "big"]
    srv = serving.Server({n: n for n in names}, {n: 4 for n in names}, capacity=64, shared=False)
    reqs = serving.poisson_requests(24, 400.0, names)
    done = serving.run_stream(srv, reqs, timeout_s=30.0)
    srv.close()
    assert len(done) == 24 and all(not q.open for q in srv.queues.values())
    by_net = {n: [r for r in reqs if r[1] == n] for n in names}
    for t in done:
        assert t.t_done - t.t_arrival >= 0.02 - 1e-3         # latency counts from arrival
        sent = by_net[t.network][t.request][2]               # request ids count per queue
        assert np.all(t.y == sent[INDEX["finaltime"]]) and t.status == 0
    assert sorted((t.network, t.request) for t in done) == sorted(
        (n, i) for n in names for i in range(len(by_net[n])))


class _FakeShared(_FakeQueue):
    """Stands in for a shared device queue: one request sequence over every network."""

    def __init__(self, nets, capacity, n_slots, lu_bits=64, idle_timeout_s=600.0):
        super().__init__(None, capacity, n_slots)
        self.nets, self.n_slots, self.which = list(nets), n_slots, []

    def submit(self, params, network):
        self.which.append(network)
        return super().submit(params)

    def result(self, request, network):
        assert self.which[request] == network
        return super().result(request)


def test_shared_server_posts_every_network_to_one_queue(monkeypatch):
    monkeypatch.setattr(serving, "SharedQueue", _FakeShared)
    names = ["small", "default", "big"]
    srv = serving.Server({n: n for n in names}, {n: 4 for n in names}, capacity=64)
    assert srv.shared_queue.n_slots == 12 and srv.shared_queue.nets == names and not srv.queues
    reqs = serving.poisson_requests(24, 400.0, names)
    done = serving.run_stream(srv, reqs, timeout_s=30.0)
    srv.close()
    assert not srv.shared_queue.open and len(done) == 24
    assert sorted(t.request for t in done) == list(range(24))      # one request sequence
    for t in done:
        arr, name, p = reqs[t.request]
        assert t.network == name and srv.shared_queue.which[t.request] == names.index(name)
        assert np.all(t.y == p[INDEX["finaltime"]])


def test_plan_shared_slots_within_cus_and_budget(fake_sizes):
    assert serving.plan_shared_slots(fake_sizes, n_cu=256, hbm_budget_bytes=1e9) == 252
    assert serving.plan_shared_slots(fake_sizes, n_cu=256, hbm_budget_bytes=14_000.0) == 10   # 1400 B per slot
    with pytest.raises(GpuError):
        serving.plan_shared_slots(fake_sizes, n_cu=256, hbm_budget_bytes=1000.0)


def test_stream_timeout_cancels_the_queues(monkeypatch):
    """A stream that does not finish in time cancels the server's queues (so that closing them
    does not wait for requests that may never end) before it reports the time-out."""
    class _Never(_FakeShared):
        def done(self, request):
            return False

    class _Net:
        def __init__(self, h):
            self.handle = h

    cancelled = []

    class _Lib:
        def ucg_cancel(self, h):
            cancelled.append(h.value)
            return 0

    monkeypatch.setattr(serving, "SharedQueue", _Never)
    monkeypatch.setattr(serving, "lib", lambda: _Lib())
    nets = {"a": _Net(3), "b": _Net(5)}
    srv = serving.Server(nets, 4, capacity=8)
    with pytest.raises(GpuError):
        serving.run_stream(srv, serving.poisson_requests(3, 1000.0, ["a", "b"]), timeout_s=0.05)
    assert cancelled == [3]          # the shared queue watches its first network's cancel word
    srv.close()
