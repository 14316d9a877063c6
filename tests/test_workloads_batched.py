"""The ensemble and serving workloads at their full sizes, and a long run checked by
invariants.

- 64 replicates (500-20,000 agents log-uniform, 2D even / 3D odd seeds, the 900-agent 2D
  workload's density and coefficients) advanced as one graph-replayed batch equal each
  replicate stepped alone, bit for bit: every field, both link lists, neighbor lists and
  voxel buckets; the mesh invariants hold per member.
- Served requests of the 88.7k-agent workload's physics (springs, attack, dead
  phagocytosis), including a 148k-agent 3D and a 180k-agent 2D request, equal the same
  requests run alone; so do requests whose step parameters differ from each other.
- 600 fused steps of the 88.7k-agent workload keep the mesh, list and count invariants.
"""
from __future__ import annotations

import numpy as np
import pytest

from paraphysicell_amd import abi, synth
from paraphysicell_amd.mechanics import Batch, DeviceSim
from paraphysicell_amd.serving import Request, Server, run_alone
from test_full_size import _check_mesh_invariants

pytestmark = pytest.mark.gpu


def _full_state(sim):
    out = {name: sim.get_field(name) for name, _, _ in abi.FIELDS}
    out["springs"] = sim.links(abi.LINKS_SPRINGS)
    out["attached"] = sim.links(abi.LINKS_ATTACHED)
    out["neighbors"] = sim.neighbors()
    out["buckets"] = sim.buckets()
    out["count"] = sim.count
    out["deaths"] = sim.deaths_count
    return out


def _assert_same(a, b, what):
    assert a["count"] == b["count"], f"{what}: count"
    assert a["deaths"] == b["deaths"], f"{what}: deaths"
    for k in a:
        if k in ("count", "deaths"):
            continue
        if isinstance(a[k], tuple):
            assert all(np.array_equal(x, y) for x, y in zip(a[k], b[k])), f"{what}: {k}"
        else:
            assert np.array_equal(a[k], b[k]), f"{what}: {k}"


def _replicate(s, n):
    mesh, st, params = synth.config3_member(s, n)
    sim = DeviceSim(mesh, 4, st.defs, st.subs, capacity=n, seed=s)
    sim.upload(st)
    sim.initialize_mesh()
    return mesh, sim, params


def test_ensemble_64_replicates_batched_equal_alone():
    sizes = synth.config3_sizes()
    steps = 12
    print(f"64-replicate ensemble: {sum(sizes)} agents, {steps} steps", flush=True)
    alone = []
    for s, n in enumerate(sizes):
        _, sim, params = _replicate(s, n)
        sim.step(params, steps)
        alone.append(_full_state(sim))
        sim.close()
    batch = Batch(4, graphs=True)
    members = []
    for s, n in enumerate(sizes):
        mesh, sim, params = _replicate(s, n)
        batch.admit(sim)
        members.append((mesh, sim))
    # graph-replayed chunks with eager reads of one member between them
    batch.step(params, 5)
    batch.synchronize()
    _ = members[3][1].get_field("positions")
    batch.step(params, steps - 5)
    batch.synchronize()
    for s, ((mesh, sim), ref) in enumerate(zip(members, alone)):
        _assert_same(_full_state(sim), ref, f"replicate {s} ({sizes[s]} agents, {mesh.dims}D)")
        _check_mesh_invariants(sim, mesh)
    batch.close()
    for _, sim in members:
        sim.close()


def _served_requests(keys, steps, arrivals, params_of=None):
    out = []
    for k, t in zip(keys, arrivals):
        mesh, st, params = synth.serving_request(k, seed=4)
        if params_of is not None:
            params = params_of(k, params)
        out.append(Request(k, float(t), mesh, st, params, steps, seed=1000 + k))
    return out


def _assert_served_equal_alone(done, alone):
    by_key = {q.key: q for q in done}
    assert len(by_key) == len(alone)
    for ref in alone:
        got = by_key[ref.key].result
        what = f"request {ref.key} ({ref.agents} agents, {ref.mesh.dims}D)"
        for name in ("count", "deaths"):
            assert got[name] == ref.result[name], f"{what}: {name}"
        for name in ("positions", "velocities", "simple_pressure", "damage"):
            assert np.array_equal(got[name], ref.result[name]), f"{what}: {name}"
        for x, y in zip(got["springs"], ref.result["springs"]):
            assert np.array_equal(x, y), f"{what}: springs"


def test_served_mix_with_large_requests_equals_alone():
    # requests 0-4 of the served workload: 148k 3D, 180k 2D, 25.6k 3D, 33.7k 2D, 4.3k 2D
    keys = [0, 1, 2, 3, 4]
    arrivals = [0.0, 0.0, 0.01, 0.02, 0.02]
    steps = 100
    alone = _served_requests(keys, steps, arrivals)
    assert max(q.agents for q in alone if q.mesh.dims == 3) >= 100_000
    assert any(q.mesh.dims == 2 for q in alone)
    print(f"served mix {[q.agents for q in alone]} agents, {steps} steps", flush=True)
    run_alone(alone)
    served = _served_requests(keys, steps, arrivals)
    server = Server(budget_bytes=200e9, chunk=10, graphs=True)
    done = server.run(served, realtime=True)
    server.close()
    _assert_served_equal_alone(done, alone)


def test_served_requests_keep_their_own_parameters():
    """Requests whose dt, wall and springs settings differ, served together, each equal the
    same request run alone under its own settings."""
    variants = {5: abi.Params(dt=0.05, virtual_wall=True, springs=True),
                6: abi.Params(dt=0.1, virtual_wall=False, springs=True),
                8: abi.Params(dt=0.1, virtual_wall=True, springs=False),
                9: abi.Params(dt=0.2, virtual_wall=False, springs=False)}

    def params_of(k, p):
        return variants.get(k, p)

    keys = [4, 5, 6, 7, 8, 9]
    arrivals = [0.0] * len(keys)
    steps = 30
    alone = _served_requests(keys, steps, arrivals, params_of)
    print(f"served with own parameters: {[q.agents for q in alone]} agents", flush=True)
    run_alone(alone)
    served = _served_requests(keys, steps, arrivals, params_of)
    server = Server(budget_bytes=200e9, chunk=7, graphs=True)
    done = server.run(served, realtime=False)
    server.close()
    _assert_served_equal_alone(done, alone)
    # the settings matter: the same request under another dt differs
    other = _served_requests([5], steps, [0.0])
    run_alone(other)
    assert not np.array_equal(other[0].result["positions"], alone[1].result["positions"])


def _symmetric(off, idx, n):
    """Whether every entry j of row i is matched by an entry i of row j, with multiplicity."""
    rows = np.repeat(np.arange(n), np.diff(off))
    a = np.stack([rows, idx], 1)
    b = np.stack([idx, rows], 1)
    a = a[np.lexsort((a[:, 1], a[:, 0]))]
    b = b[np.lexsort((b[:, 1], b[:, 0]))]
    return np.array_equal(a, b)


def test_workload_88k_600_steps_invariants():
    mesh, st, params = synth.config2(4)
    n0 = st.count
    sim = DeviceSim(mesh, 4, st.defs, st.subs, capacity=n0, seed=2)
    sim.upload(st)
    sim.initialize_mesh()
    print("88.7k-agent workload, 600 steps", flush=True)
    sim.step(params, 600)
    sim.synchronize()
    n = sim.count
    deaths = sim.deaths_count
    assert 0 < n <= n0
    # every removal is a death (flagged) or an agent that left the box; none left the box
    # here (the ball stays far from the walls), so removals are the deaths
    assert n0 - n == deaths
    pos = sim.get_field("positions").reshape(n, 3)
    assert np.all(np.isfinite(pos))
    assert np.all((pos >= np.array(mesh.bbox_min)) & (pos <= np.array(mesh.bbox_max)))
    for name in ("velocities", "previous_velocities", "simple_pressure", "damage"):
        assert np.all(np.isfinite(sim.get_field(name))), name
    _check_mesh_invariants(sim, mesh)
    # springs: symmetric lists within capacity; never self-attached
    soff, sidx = sim.links(abi.LINKS_SPRINGS)
    assert np.all(np.diff(soff) <= 12)
    assert _symmetric(soff, sidx, n)
    assert not np.any(sidx == np.repeat(np.arange(n), np.diff(soff)))
    # neighbor lists equal the brute-force rule on the final state (sets, movable agents)
    off, idx = sim.neighbors()
    radius = sim.get_field("radius").astype(np.float32)
    rel = sim.get_field("relative_maximum_adhesion_distance").astype(np.float32)
    movable = sim.get_field("is_movable").astype(bool)
    reach = (rel * radius).astype(np.float32)
    g = np.array(mesh.grid_shape())
    vox = np.clip(((pos - np.array(mesh.bbox_min, np.float32)) / np.float32(30)).astype(np.int64), 0, g - 1)
    rng = np.random.default_rng(0)
    sample = rng.choice(np.nonzero(movable)[0], size=2000, replace=False)
    for i in sample:
        cand = np.nonzero(np.all(np.abs(vox - vox[i]) <= 1, axis=1))[0]
        cand = cand[cand != i]
        d = pos[cand] - pos[i]
        s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        within = np.sqrt(s, dtype=np.float32) <= reach[i] + reach[cand]
        assert set(cand[within].tolist()) == set(idx[off[i]:off[i + 1]].tolist()), f"neighbors of {i}"
    # only movable agents have lists; between movable agents the rule is symmetric
    rows = np.repeat(np.arange(n), np.diff(off))
    assert np.all(movable[rows])
    keep = movable[idx]
    pairs = np.stack([rows[keep], idx[keep]], 1)
    flipped = pairs[:, ::-1]
    order = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))]  # noqa: E731
    assert np.array_equal(order(pairs), order(flipped))
    sim.close()
