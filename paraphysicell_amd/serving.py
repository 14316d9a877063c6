"""Ensembles and a served request loop over batches (pcm_batch_*).

``run_ensemble`` advances a list of simulations together (one segmented launch per kernel
for all members of a dimensionality) and, for comparison, one at a time.

``Server`` is a continuous-batching request loop.  Each request is one new simulation and
a number of mechanics steps.  A request that has arrived is admitted as soon as the HBM
budget allows: its state is uploaded, its mesh built, and it joins the live batch between
two step chunks, while the other requests keep advancing.  A finished request leaves the
batch, its result is read back, and its device memory is freed.  Latency runs from arrival
to result.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .mechanics import Batch, DeviceSim


def footprint_bytes(mesh: abi.MeshSpec, real_bytes: int, defs: int, subs: int, capacity: int) -> int:
    """Device bytes of one simulation: agent fields, fixed-stride lists and step scratch per
    agent, plus two ints per mechanics voxel."""
    d, r = mesh.dims, real_bytes
    fields = sum(abi.field_width(w, d, defs, subs) * (r if t == abi.T_REAL else (4 if t == abi.T_I32 else 1))
                 for _, t, w in abi.FIELDS)
    k = {1: 8, 2: 16, 3: 32}[d]
    # voxel tables (a range of 32 bucket slots per occupied voxel, 32 table slots per voxel):
    # 3D, and 1D and 2D below the lane-per-agent neighbor search (capacity under 4096)
    tables = d == 3 or capacity < 4096
    b = 32 if tables else 1  # bucket slots per agent
    per_agent = fields + 4 * (k + 12 + 1 + 2 * k + k) + k + 12 * r + 160 + (4 + 4 * r) * (b - 1)
    g = mesh.grid_shape()
    per_voxel = 8 + (4 * b if tables else 0)  # counts, starts and the unordered tables
    return int(capacity * per_agent + per_voxel * g[0] * g[1] * g[2] + (1 << 16))


@dataclass
class Request:
    key: int
    arrival: float  # seconds after the start of the run
    mesh: abi.MeshSpec
    state: object
    params: abi.Params
    steps: int
    seed: int = 0
    # filled by the server
    sim: DeviceSim | None = None
    remaining: int = 0
    need: int = 0
    admitted: float = float("nan")
    finished: float = float("nan")
    result: dict = field(default_factory=dict)

    @property
    def agents(self) -> int:
        return self.state.count

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _start(req: Request) -> DeviceSim:
    st = req.state
    sim = DeviceSim(req.mesh, st.real_bytes, st.defs, st.subs, capacity=st.count, seed=req.seed)
    sim.upload(st)
    sim.initialize_mesh()
    sim.set_step_params(req.params)  # the request's own dt, wall and springs in the batch
    return sim


def _result(sim: DeviceSim) -> dict:
    out = sim.get_fields(("positions", "velocities", "simple_pressure", "damage"))
    out.update(count=sim.count, deaths=sim.deaths_count, springs=sim.links(abi.LINKS_SPRINGS))
    return out


def _finish(sim: DeviceSim):
    """Read a finished request's result back and free its simulation (memory returns to the
    library's pool); returns the result and the time it was ready."""
    result = _result(sim)
    sim.close()
    return result, time.perf_counter()


class Server:
    """Continuous batching over one pcm batch (hipGraph-captured steps by default).

    The stepping thread only advances the batch and moves requests in and out of it; set-up
    (handle, upload, first mesh and neighbor build) and tear-down (result read-back, free)
    run on worker threads, each on the request's own stream, so they overlap the batch's
    steps on the device.  A request counts against the HBM budget from the start of its
    set-up to the end of its tear-down.

    `tuning` maps a dimensionality (2 or 3) to a launch shape (velocity lanes, event lanes,
    neighbor waves) for that group's launches, e.g. one that pcm_batch_autotune picked on
    the same kind of workload; requests are not tuned on while they are served."""

    def __init__(self, budget_bytes: float = 200e9, chunk: int = 10, graphs: bool = True, workers: int = 8,
                 tuning: dict | None = None):
        from concurrent.futures import ThreadPoolExecutor

        self.budget = budget_bytes
        self.chunk = chunk
        self.batch = Batch(4, graphs=graphs)
        for dims, shape in (tuning or {}).items():
            self.batch.set_tuning(dims, shape)
        self.used = 0
        self.agent_steps = 0
        self.workers = ThreadPoolExecutor(max_workers=max(1, workers), thread_name_prefix="pcm-serve")

    @staticmethod
    def order(q: Request):
        """Set-up order: arrival, then the largest of simultaneous arrivals first."""
        return (q.arrival, -q.agents, q.key)

    def run(self, requests: list[Request], realtime: bool = True) -> list[Request]:
        """Serve `requests`.  With realtime, a request only becomes visible at its arrival
        time; without, all are present at once (saturation).

        Requests start their set-up in arrival order, the largest first among those that
        arrive together (each request takes the same number of steps, so a large one that
        joined last would step on alone at the end).  The batch steps in chunks whenever it
        has members; finished set-ups join, and finished requests leave, between chunks."""
        from concurrent.futures import FIRST_COMPLETED, wait

        pending = sorted(requests, key=self.order)
        starting: list[tuple[Request, object]] = []   # set-up in flight
        live: list[Request] = []                       # in the batch
        finishing: list[tuple[Request, object]] = []  # tear-down in flight
        counts: dict[int, int] = {}                    # agents of each live request (by key)
        done: list[Request] = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        while pending or starting or live or finishing:
            # arrived requests start their set-up while the budget allows
            while pending and (not realtime or pending[0].arrival <= now()):
                q = pending[0]
                need = footprint_bytes(q.mesh, q.state.real_bytes, q.state.defs, q.state.subs, q.agents)
                if (live or starting or finishing) and self.used + need > self.budget:
                    break
                pending.pop(0)
                q.need = need
                q.admitted = now()
                self.used += need
                starting.append((q, self.workers.submit(_start, q)))
            # finished set-ups join the batch between two step chunks
            for item in [x for x in starting if x[1].done()]:
                q, fut = item
                starting.remove(item)
                q.sim = fut.result()
                q.remaining = q.steps
                self.batch.admit(q.sim)
                counts[q.key] = q.agents
                live.append(q)
            for item in [x for x in finishing if x[1].done()]:
                q, fut = item
                finishing.remove(item)
                q.result, ready = fut.result()
                q.finished = ready - t0
                self.used -= q.need
                done.append(q)
            if not live:
                # nothing to step: wait for a set-up or tear-down to end, or the next arrival
                futures = [f for _, f in starting] + [f for _, f in finishing]
                limit = 0.05
                if pending and realtime:
                    limit = min(limit, max(0.0, pending[0].arrival - now()))
                if futures:
                    wait(futures, timeout=limit, return_when=FIRST_COMPLETED)
                elif pending:
                    time.sleep(min(limit, 0.002))
                continue
            k = min(self.chunk, min(q.remaining for q in live))
            before = sum(counts[q.key] for q in live)
            self.batch.step(None, k)  # every member under its own request's parameters
            after_counts = self.batch.counts([q.sim for q in live])
            for q, n in zip(live, after_counts):
                counts[q.key] = n
            self.agent_steps += 0.5 * (before + sum(after_counts)) * k
            for q in list(live):
                q.remaining -= k
                if q.remaining == 0:
                    self.batch.retire(q.sim)
                    live.remove(q)
                    del counts[q.key]
                    finishing.append((q, self.workers.submit(_finish, q.sim)))
                    q.sim = None
        self.elapsed = now()
        return done

    def close(self):
        self.workers.shutdown(wait=True)
        self.batch.close()


def run_alone(requests: list[Request], shapes: list | None = None) -> tuple[float, int]:
    """Each request on its own, one after another: (seconds, agent-steps).  `shapes`: the
    launch shape of each request (as the tuner picked it for that simulation alone), or None
    for the library default."""
    t0 = time.perf_counter()
    agent_steps = 0
    for k, q in enumerate(requests):
        sim = _start(q)
        if shapes is not None:
            sim.tuning = shapes[k]
        before = sim.count
        sim.step(q.params, q.steps)
        agent_steps += 0.5 * (before + sim.count) * q.steps
        q.result = _result(sim)
        sim.close()
    return time.perf_counter() - t0, agent_steps


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times of a Poisson stream (exponential gaps, seeded)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))


def run_ensemble(members, params: abi.Params, steps: int, batched: bool = True, graphs: bool = True,
                 autotune: bool = True, shapes: dict | None = None):
    """Advance every (mesh, state, seed) member `steps` steps; returns (seconds, agent-steps,
    sims).  The timed region covers the steps only: states are resident and the launch
    shape tuned (batch groups, or each simulation alone) before it starts.  The shapes used
    go into `shapes`: dims -> shape of that group when batched, member index -> its shape
    when one at a time."""
    sims = []
    for mesh, st, seed in members:
        sim = DeviceSim(mesh, st.real_bytes, st.defs, st.subs, capacity=st.count, seed=seed)
        sim.upload(st)
        sim.initialize_mesh()
        sims.append(sim)
    agents = sum(s.count for s in sims)
    batch = None
    if batched:
        batch = Batch(4, graphs=graphs)
        for s in sims:
            batch.admit(s)
        if autotune:
            batch.autotune(params, 2)  # launch shape per dims group, on real steps
        if shapes is not None:
            for d in sorted({s.mesh.dims for s in sims}):
                shapes[d] = batch.tuning(d)
        batch.step(params, 1)  # warm-up: first-step refresh and graph capture
        batch.synchronize()
    else:
        for k, s in enumerate(sims):
            if autotune:
                s.autotune(params, 2)
            if shapes is not None:
                shapes[k] = s.tuning
            s.step(params, 1)
            s.synchronize()
    t0 = time.perf_counter()
    if batched:
        batch.step(params, steps)
        batch.synchronize()
    else:
        for s in sims:
            s.step(params, steps)
        for s in sims:
            s.synchronize()
    elapsed = time.perf_counter() - t0
    agent_steps = sum(s.count for s in sims) * 0.5 * steps + agents * 0.5 * steps
    if batch is not None:
        batch.close()
    return elapsed, agent_steps, sims
