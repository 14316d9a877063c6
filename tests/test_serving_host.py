"""Host logic of the served request loop (paraphysicell_amd/serving.py), on the CPU: the
device batch and simulations are replaced by fakes that count steps, so the admission by
HBM budget, the chunked stepping, the set-up order and the agent-step accounting are
checked without a GPU."""
from __future__ import annotations

import threading
import time
from types import SimpleNamespace

import pytest

from paraphysicell_amd import abi, serving
from paraphysicell_amd.serving import Request, Server, footprint_bytes


class FakeSim:
    """A simulation that loses one agent every `loss_every` steps."""

    def __init__(self, q, loss_every):
        self.key = q.key
        self.count = q.agents
        self.steps = 0
        self.loss_every = loss_every
        self._h = object()

    def advance(self, k):
        for _ in range(k):
            self.steps += 1
            if self.loss_every and self.steps % self.loss_every == 0:
                self.count -= 1


class FakeBatch:
    def __init__(self, real_bytes=4, graphs=False, lib=None):
        self.members = []
        self.chunks = []  # (k, member keys)
        self.tuning_set = {}

    def admit(self, sim):
        assert sim not in self.members
        self.members.append(sim)

    def retire(self, sim):
        self.members.remove(sim)

    def step(self, params, k):
        assert params is None  # every member steps under its own parameters
        self.chunks.append((k, sorted(s.key for s in self.members)))
        for s in self.members:
            s.advance(k)

    def counts(self, sims):
        for s in sims:
            assert s in self.members
        return [s.count for s in sims]

    def set_tuning(self, dims, shape):
        self.tuning_set[dims] = tuple(shape)

    def close(self):
        self.members = []


class Harness:
    """Fake set-up and tear-down that record the order and the device memory in use."""

    def __init__(self, loss_every=0):
        self.loss_every = loss_every
        self.lock = threading.Lock()
        self.started = []
        self.finished = []
        self.sims = {}

    def start(self, q):
        with self.lock:
            self.started.append(q.key)
        time.sleep(0.001)
        sim = FakeSim(q, self.loss_every)
        self.sims[q.key] = sim
        return sim

    def finish(self, sim):
        with self.lock:
            self.finished.append(sim.key)
        return {"count": sim.count, "steps": sim.steps}, time.perf_counter()


def _requests(sizes, arrivals, steps):
    out = []
    for k, (n, t) in enumerate(zip(sizes, arrivals)):
        mesh = abi.MeshSpec(3, (-200,) * 3, (200,) * 3)
        st = SimpleNamespace(count=n, real_bytes=4, defs=2, subs=1)
        out.append(Request(k, t, mesh, st, abi.Params(), steps))
    return out


@pytest.fixture
def fakes(monkeypatch):
    h = Harness()
    monkeypatch.setattr(serving, "Batch", FakeBatch)
    monkeypatch.setattr(serving, "_start", h.start)
    monkeypatch.setattr(serving, "_finish", h.finish)
    return h


def test_every_request_takes_exactly_its_steps(fakes):
    sizes = [500, 4000, 120, 2500, 900, 60, 3000]
    reqs = _requests(sizes, [0.0] * len(sizes), steps=23)
    server = Server(budget_bytes=1e12, chunk=5, tuning={3: (4, 4, 2)})
    done = server.run(reqs, realtime=False)
    assert sorted(q.key for q in done) == list(range(len(sizes)))
    for q in done:
        assert q.result["steps"] == 23, q.key
        assert q.latency > 0
    assert server.batch.tuning_set == {3: (4, 4, 2)}
    # chunks never run a member past its steps, and members join and leave between chunks
    assert all(k <= 5 for k, _ in server.batch.chunks)
    server.close()


def test_largest_simultaneous_arrival_is_set_up_first(fakes):
    sizes = [500, 4000, 120, 2500, 900]
    reqs = _requests(sizes, [0.0, 0.0, 0.0, 0.0, 0.0], steps=3)
    server = Server(budget_bytes=1e12, chunk=10, workers=1)
    server.run(reqs, realtime=False)
    server.close()
    assert fakes.started == [1, 3, 4, 0, 2]  # by size, largest first


def test_agent_steps_follow_the_counts(monkeypatch):
    h = Harness(loss_every=4)
    monkeypatch.setattr(serving, "Batch", FakeBatch)
    monkeypatch.setattr(serving, "_start", h.start)
    monkeypatch.setattr(serving, "_finish", h.finish)
    sizes = [100, 40]
    reqs = _requests(sizes, [0.0, 0.0], steps=8)
    server = Server(budget_bytes=1e12, chunk=4, workers=1)
    done = server.run(reqs, realtime=False)
    server.close()
    # each member loses an agent every 4 steps: the mean of the counts before and after each
    # chunk of 4 steps, times 4
    expected = 0.0
    for n in sizes:
        per_member = 0.0
        count = n
        for _ in range(2):
            per_member += 0.5 * (count + count - 1) * 4
            count -= 1
        expected += per_member
    assert server.agent_steps == pytest.approx(expected)
    assert {q.key: q.result["count"] for q in done} == {0: 98, 1: 38}


def test_admission_keeps_to_the_budget(fakes):
    sizes = [3000, 2000, 1500, 1000, 800, 600]
    reqs = _requests(sizes, [0.0] * len(sizes), steps=6)
    need = [footprint_bytes(q.mesh, 4, 2, 1, q.agents) for q in reqs]
    budget = need[0] + need[1]  # room for the two largest together
    server = Server(budget_bytes=budget, chunk=2, workers=2)
    in_use = []
    orig_start = fakes.start

    def start(q):
        in_use.append(server.used)
        return orig_start(q)

    serving._start = start
    done = server.run(reqs, realtime=False)
    server.close()
    assert len(done) == len(sizes)
    assert max(in_use) <= budget
    assert server.used == 0  # every request's memory returned


def test_requests_are_not_admitted_before_they_arrive(fakes):
    sizes = [200, 300, 400]
    arrivals = [0.0, 0.03, 0.06]
    reqs = _requests(sizes, arrivals, steps=2)
    server = Server(budget_bytes=1e12, chunk=2, workers=2)
    done = server.run(reqs, realtime=True)
    server.close()
    for q in done:
        assert q.admitted >= q.arrival
        assert q.finished >= q.admitted
