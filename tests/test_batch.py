"""Batches (pcm_batch_*): simulations of different sizes and dimensionalities advanced
together give the same results as each simulation stepped alone, also when members
join and leave between steps and when a member's neighbor capacity overflows."""
from __future__ import annotations

import numpy as np
import pytest

from paraphysicell_amd import abi
from paraphysicell_amd.mechanics import Batch, DeviceSim
from parity import rich_population

pytestmark = pytest.mark.gpu


def _state_of(sim):
    out = {name: sim.get_field(name) for name, _, _ in abi.FIELDS}
    out["springs"] = sim.links(abi.LINKS_SPRINGS)
    out["attached"] = sim.links(abi.LINKS_ATTACHED)
    out["neighbors"] = sim.neighbors()
    out["count"] = sim.count
    out["deaths"] = sim.deaths_count
    return out


def _assert_equal(a, b, what):
    assert a["count"] == b["count"], what
    assert a["deaths"] == b["deaths"], what
    for k in a:
        if k in ("count", "deaths"):
            continue
        if isinstance(a[k], tuple):
            assert all(np.array_equal(x, y) for x, y in zip(a[k], b[k])), f"{what}: {k}"
        else:
            assert np.array_equal(a[k], b[k]), f"{what}: {k}"


def _member(spec, real_bytes=4):
    dims, n, seed = spec
    mesh, st, params = rich_population(dims, n, seed, real_bytes)
    sim = DeviceSim(mesh, real_bytes, st.defs, st.subs, capacity=n, seed=seed + 1000)
    sim.upload(st)
    sim.initialize_mesh()
    return sim, params


SPECS = [(2, 40, 1), (3, 300, 2), (2, 500, 3), (3, 64, 4), (3, 900, 5)]


@pytest.mark.parametrize("flatten", [True, False])
@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("real_bytes", [4, 8])
def test_batch_equals_alone(real_bytes, graphs, flatten):
    print(f"batch vs alone, real_bytes {real_bytes}, graphs {graphs}, flattened grids {flatten}", flush=True)
    alone = []
    for spec in SPECS:
        sim, params = _member(spec, real_bytes)
        sim.step(params, 12)
        alone.append(_state_of(sim))
        sim.close()
    batch = Batch(real_bytes, graphs=graphs)
    batch.use_flattened_grids(flatten)
    sims = []
    for spec in SPECS:
        sim, params = _member(spec, real_bytes)
        batch.admit(sim)
        sims.append(sim)
    assert batch.size == len(SPECS)
    batch.step(params, 5)
    # eager work on a member between replays (a read of its state)
    _ = sims[1].get_field("positions")
    batch.step(params, 7)
    batch.synchronize()
    for spec, sim, ref in zip(SPECS, sims, alone):
        _assert_equal(_state_of(sim), ref, f"member {spec}")
    batch.close()


def test_batch_graph_replays_match_alone_each_chunk():
    """Every graph replay chunk agrees with the simulations stepped alone, with eager reads
    and synchronisation between the replays."""
    print("graph replays vs alone", flush=True)
    specs = SPECS[:3]
    refs = []
    for spec in specs:
        sim, params = _member(spec)
        refs.append(sim)
    batch = Batch(4, graphs=True)
    sims = []
    for spec in specs:
        sim, params = _member(spec)
        batch.admit(sim)
        sims.append(sim)
    for chunk in (3, 1, 4, 2):
        batch.step(params, chunk)
        batch.synchronize()
        for spec, sim, ref in zip(specs, sims, refs):
            ref.step(params, chunk)
            _assert_equal(_state_of(sim), _state_of(ref), f"member {spec} after chunk {chunk}")
    batch.close()


def test_batch_join_and_leave():
    """A joins at step 0, B after 4 steps, A leaves after 8 and steps 3 more alone."""
    print("batch join and leave", flush=True)
    refs = {}
    for name, spec, steps in (("A", (3, 400, 11), 11), ("B", (2, 300, 12), 9)):
        sim, params = _member(spec)
        sim.step(params, steps)
        refs[name] = _state_of(sim)
        sim.close()
    a, params = _member((3, 400, 11))
    b, _ = _member((2, 300, 12))
    batch = Batch(4, graphs=True)
    batch.admit(a)
    batch.step(params, 4)
    batch.admit(b)
    batch.step(params, 4)
    batch.retire(a)
    a.step(params, 3)
    batch.step(params, 5)
    batch.synchronize()
    _assert_equal(_state_of(a), refs["A"], "A")
    _assert_equal(_state_of(b), refs["B"], "B")
    batch.close()


def test_batch_member_neighbor_overflow():
    """A member whose neighbor lists outgrow their capacity recovers inside the batch."""
    print("batch neighbor overflow", flush=True)

    def dense():
        mesh = abi.MeshSpec(3, (-100,) * 3, (100,) * 3)
        rng = np.random.default_rng(21)
        from paraphysicell_amd.state import AgentState

        st = AgentState(mesh, 4, 1, 1, 120).defaults()
        st["positions"] = rng.uniform(-6, 6, 360)  # far more than 32 neighbors each
        sim = DeviceSim(mesh, 4, 1, 1, capacity=120, seed=3)
        sim.upload(st)
        return sim, abi.Params(dt=0.1, virtual_wall=True, springs=False)

    # no initial mesh: the first step's neighbor build overflows the initial capacity
    ref, params = dense()
    ref.step(params, 3)
    want = _state_of(ref)
    ref.close()
    other_ref, oparams = _member((3, 200, 8))
    other_ref.step(params, 3)
    other_want = _state_of(other_ref)
    other_ref.close()

    sim, _ = dense()
    other, _ = _member((3, 200, 8))
    batch = Batch(4)
    batch.admit(other)
    batch.admit(sim)
    batch.step(params, 3)
    batch.synchronize()
    _assert_equal(_state_of(sim), want, "dense member")
    _assert_equal(_state_of(other), other_want, "other member")
    batch.close()


SHAPES = [(4, 4, 1), (8, 8, 2), (16, 16, 4), (4, 16, 2), (16, 8, 1)]


@pytest.mark.parametrize("spec", [(3, 400, 31), (2, 300, 33)])
@pytest.mark.parametrize("real_bytes", [4, 8])
def test_launch_shapes_give_identical_results(real_bytes, spec):
    """Every launch shape of the lane-parallel kernels gives bit-identical states."""
    print(f"launch shapes, real_bytes {real_bytes}, {spec}", flush=True)
    want = None
    for shape in SHAPES:
        sim, params = _member(spec, real_bytes)
        sim.tuning = shape
        assert sim.tuning == shape
        sim.step(params, 6)
        got = _state_of(sim)
        sim.close()
        if want is None:
            want = got
        else:
            _assert_equal(got, want, f"shape {shape}")


def test_autotune_keeps_results():
    """Tuning on the simulation's own steps changes nothing but the launch shape."""
    print("autotune", flush=True)
    ref, params = _member((3, 600, 32))
    ref.step(params, 1 + 6 * 2 + 4)
    sim, _ = _member((3, 600, 32))
    chosen = sim.autotune(params, 2)
    assert sim.tuning == chosen
    sim.step(params, 4)
    _assert_equal(_state_of(sim), _state_of(ref), "autotuned")
    batch = Batch(4, graphs=True)
    refs = []
    sims = []
    for spec in SPECS[:3]:
        r, _ = _member(spec)
        r.step(params, 1 + 6 * 2 + 3)
        refs.append(r)
        s, _ = _member(spec)
        batch.admit(s)
        sims.append(s)
    batch.autotune(params, 2)
    batch.step(params, 3)
    batch.synchronize()
    for spec, s, r in zip(SPECS[:3], sims, refs):
        _assert_equal(_state_of(s), _state_of(r), f"batch autotuned member {spec}")
    batch.close()


def test_batch_counts_shapes_and_bulk_reads():
    """pcm_batch_counts gives each member's count as the last step left it (the scalars the
    step gathered), pcm_batch_set_tuning's shape is the one the group runs and keeps for a
    member that joins later, and pcm_get_fields reads what pcm_get_field reads."""
    print("batch counts, group shapes, bulk field reads", flush=True)
    sims = [_member(spec) for spec in SPECS]
    params = sims[0][1]
    batch = Batch(4, graphs=True)
    batch.set_tuning(3, (4, 16, 1))
    assert batch.tuning(3) == (4, 16, 1)
    for sim, _ in sims[:3]:
        batch.admit(sim)
    batch.step(params, 6)
    assert batch.counts([s for s, _ in sims[:3]]) == [s.count for s, _ in sims[:3]]
    for sim, _ in sims[3:]:
        batch.admit(sim)
    batch.step(params, 6)
    assert batch.tuning(3) == (4, 16, 1)
    members = [s for s, _ in sims]
    assert batch.counts(members) == [s.count for s in members]
    with pytest.raises(Exception):
        batch.counts([DeviceSim(*_member_args(SPECS[0]))])
    names = ("positions", "velocities", "simple_pressure", "damage", "cell_definition_index", "dead")
    for sim in members:
        bulk = sim.get_fields(names)
        for name in names:
            assert np.array_equal(bulk[name], sim.get_field(name)), name
    batch.close()
    for sim in members:
        sim.close()


def _member_args(spec):
    dims, n, seed = spec
    mesh, st, _ = rich_population(dims, n, seed, 4)
    return mesh, 4, st.defs, st.subs, n, seed
