"""Parity of the MI355X library (through the C-ABI) with the oracle and the reference tests.

Bar: voxel buckets and neighbor lists bit-exact; integer/byte fields and link lists
bit-exact; real fields within 1e-5 (FP32) / 1e-12 (FP64) relative to max(|oracle|, 1).
"""
import dataclasses

import numpy as np
import pytest

import parity
import scenarios as sc
from paraphysicell_amd import abi
from paraphysicell_amd.mechanics import DeviceSim
from test_oracle_reference_cases import float_eq

pytestmark = pytest.mark.gpu


def device(mesh, st, seed=7, capacity=None):
    d = DeviceSim(mesh, st.real_bytes, st.defs, st.subs, capacity=capacity or max(st.count, 1), seed=seed)
    d.upload(st)
    return d


# ------------------------------------------------------------------ reference test cases
@pytest.mark.parametrize("real_bytes", [4, 8])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_velocity_simple(dims, real_bytes):
    mesh, st, p = sc.velocity_simple(dims, real_bytes)
    d = device(mesh, st)
    d.update_mechanics_mesh()
    d.update_cell_neighbors()
    d.update_cell_forces(p)
    v = d.get_field("velocities").reshape(3, dims)
    expected = sc.VELOCITY_SIMPLE_EXPECTED[dims]
    for k in range(dims):
        assert float_eq(v[0, k], expected) and float_eq(v[1, k], -expected)
        assert v[2, k] == 0


@pytest.mark.parametrize("real_bytes", [4, 8])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_velocity_complex(dims, real_bytes):
    mesh, st, p = sc.velocity_complex(dims, real_bytes)
    d = device(mesh, st)
    d.update_mechanics_mesh()
    d.update_cell_neighbors()
    d.update_cell_forces(p)
    off, _ = d.neighbors()
    assert list(np.diff(off)) == [1, 2, 1]
    v = d.get_field("velocities").reshape(3, dims)
    expected = sc.brute_force_velocities(st)
    for c in range(3):
        for k in range(dims):
            assert float_eq(v[c, k], expected[c, k])


@pytest.mark.parametrize("adhesion", [1, 3])
@pytest.mark.parametrize("per_voxel", [1, 5, 50])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_neighbors_grid(dims, per_voxel, adhesion):
    mesh, st, _ = sc.neighbors_grid(dims, per_voxel, adhesion)
    dev, ora = parity.make_pair(mesh, st)
    for s in (dev, ora):
        s.update_mechanics_mesh()
    counts, _ = dev.buckets()
    assert np.all(counts == per_voxel)
    parity.assert_same_buckets(dev, ora, "grid buckets")
    for s in (dev, ora):
        s.update_cell_neighbors()
    parity.assert_same_neighbors(dev, ora, "grid neighbors")
    off, _ = dev.neighbors()
    grid = mesh.grid_shape()
    pos = st.view("positions")
    for i in range(0, st.count, max(1, st.count // 200)):
        c = [int((pos[i, a] - mesh.bbox_min[a]) // mesh.voxel[a]) if a < dims else 0 for a in range(3)]
        assert off[i + 1] - off[i] == sc.expected_grid_neighbor_count(c, grid, per_voxel, adhesion)


@pytest.mark.parametrize("kind", ["attack", "dead_phagocytosis", "phagocytosis", "fusion"])
def test_interactions_reference(kind):
    mesh, st, p = sc.interactions_case(kind, 4)
    dev, ora = parity.make_pair(mesh, st)
    for s in (dev, ora):
        s.update_mechanics_mesh()
        s.update_cell_neighbors()
        s.update_cell_cell_interactions(p)
        if kind in ("phagocytosis", "fusion"):
            s.update_cell_cell_interactions(p)
    flags = list(dev.get_field("flags"))
    if kind == "attack":
        dmg = dev.get_field("damage")
        assert dmg[0] == 0 and dmg[1] == 0 and dmg[2] != 0
    else:
        assert flags == [abi.FLAG_NONE, abi.FLAG_TO_REMOVE, abi.FLAG_NONE]
    parity.assert_same_state(dev, ora, what=kind)


def test_container_add_and_remove():
    mesh = abi.MeshSpec(1, (0, 0, 0), (80, 0, 0))
    st = sc.zero_cell_state(mesh, 4, 3, 2, 4)
    st["volume"] = [1, 2, 3, 4]
    st["velocities"] = [10, 20, 30, 40]
    st["positions"] = [10, 20, 30, 40]
    st["flags"][2] = abi.FLAG_TO_REMOVE
    d = device(mesh, st)
    d.update_cell_container_for_mechanics(abi.Params())
    assert list(d.get_field("volume")) == [1, 2, 4]
    assert list(d.get_field("velocities")) == [10, 20, 40]
    s2 = d.download()
    s2["flags"][0] = abi.FLAG_TO_REMOVE
    d.upload(s2)
    d.update_cell_container_for_mechanics(abi.Params())
    assert list(d.get_field("volume")) == [4, 2]
    assert list(d.get_field("velocities")) == [40, 20]
    assert d.deaths_count == 2


# ------------------------------------------------------------------ per-phase parity
PHASES = ["evaluate_interactions", "update_cell_forces", "update_motility", "update_basement_membrane_interactions",
          "update_spring_attachments", "update_positions", "update_cell_cell_interactions",
          "update_cell_container_for_mechanics", "update_mechanics_mesh", "update_cell_neighbors"]


@pytest.mark.parametrize("rates", ["busy", "rare"])
@pytest.mark.parametrize("real_bytes", [4, 8])
@pytest.mark.parametrize("dims", [2, 3])
def test_phase_by_phase(dims, real_bytes, rates):
    mesh, st, p = parity.rich_population(dims, 1500 if dims == 3 else 800, 10 + dims, real_bytes, rates=rates)
    dev, ora = parity.make_pair(mesh, st, seed=5)
    for s in (dev, ora):
        s.initialize_mesh()
    parity.assert_same_buckets(dev, ora, "initial mesh")
    parity.assert_same_neighbors(dev, ora, "initial neighbors")
    for step in range(3):
        for ph in PHASES:
            getattr(dev, ph)(p)
            getattr(ora, ph)(p)
            what = f"step {step} after {ph}"
            parity.assert_same_state(dev, ora, what=what)
            if ph == "update_mechanics_mesh":
                parity.assert_same_buckets(dev, ora, what)
            if ph == "update_cell_neighbors":
                parity.assert_same_neighbors(dev, ora, what)
        # continue both from the oracle's state so rounding differences never accumulate
        dev.upload(ora.download())
        dev.step_index = ora.step_index
        dev.initialize_mesh()


@pytest.mark.parametrize("real_bytes", [4, 8])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_fused_step(dims, real_bytes):
    """pcm_step (all phases fused into one enqueue) against the oracle's step."""
    n = {1: 200, 2: 800, 3: 2000}[dims]
    mesh, st, p = parity.rich_population(dims, n, 20 + dims, real_bytes, rates="busy")
    dev, ora = parity.make_pair(mesh, st, seed=9)
    for s in (dev, ora):
        s.initialize_mesh()
    rounds, replays = [], []
    for step in range(4):
        d0, o0 = dev.deaths_count, ora.deaths_count
        dev.step(p, 1)
        ora.step(p, 1)
        assert dev.deaths_count - d0 == ora.deaths_count - o0, f"fused step {step}: deaths"
        parity.assert_same_state(dev, ora, what=f"fused step {step}")
        parity.assert_same_buckets(dev, ora, f"fused step {step}")
        parity.assert_same_neighbors(dev, ora, f"fused step {step}")
        stats = dev.stats()
        rounds.append(stats["resolve_rounds"])
        replays.append(stats["interaction_replay"])
        dev.upload(ora.download())
        dev.step_index = ora.step_index
        dev.initialize_mesh()
    if dims > 1:
        # the busy rates leave attachment events undecided after the first resolver round,
        # so the later rounds (the single-block resolver) are covered too
        assert max(rounds) >= 2, rounds
    # and their interaction events collide, so the exact serial replay ran and was compared
    assert max(replays) > 0, replays


@pytest.mark.parametrize("real_bytes", [4, 8])
def test_fused_step_with_mass_removal(real_bytes):
    """A busy 20k-agent population loses thousands of agents per step to ingestion, fusion
    and the wall: the interaction replay then touches more agents than its LDS list holds,
    so the container plan rescans the whole population (smaller populations settle only the
    touched agents, covered by test_fused_step)."""
    mesh, st, p = parity.rich_population(3, 20000, 77, real_bytes, box=323, rates="busy")
    dev, ora = parity.make_pair(mesh, st, seed=5)
    for s in (dev, ora):
        s.initialize_mesh()
    for step in range(3):
        print(f"mass removal step {step} real_bytes {real_bytes}", flush=True)
        d0, o0 = dev.deaths_count, ora.deaths_count
        dev.step(p, 1)
        ora.step(p, 1)
        deaths = ora.deaths_count - o0
        assert dev.deaths_count - d0 == deaths, f"mass removal step {step}: deaths"
        parity.assert_same_state(dev, ora, what=f"mass removal step {step}")
        parity.assert_same_buckets(dev, ora, f"mass removal step {step}")
        parity.assert_same_neighbors(dev, ora, f"mass removal step {step}")
        assert dev.stats()["interaction_replay"] > 0
        if step == 0:
            # well past 512 ingestions and fusions (two touched agents each)
            assert deaths > 1500, deaths
        dev.upload(ora.download())
        dev.step_index = ora.step_index
        dev.initialize_mesh()


@pytest.mark.parametrize("real_bytes", [4, 8])
def test_consecutive_fused_steps(real_bytes):
    """Fused steps back to back, with no upload between them, against the oracle: from the
    second step on, the detachment masks come from the previous step's mesh update instead
    of their own launch; a parameter change in between brings that launch back."""
    mesh, st, p = parity.rich_population(3, 1500, 31, real_bytes, rates="busy")
    dev, ora = parity.make_pair(mesh, st, seed=4)
    for s in (dev, ora):
        s.initialize_mesh()
    dev.step(p, 3)
    ora.step(p, 3)
    assert dev.deaths_count == ora.deaths_count, "three fused steps: deaths"
    parity.assert_same_state(dev, ora, what="three fused steps")
    parity.assert_same_neighbors(dev, ora, "three fused steps")
    q = dataclasses.replace(p, dt=0.5 * p.dt)
    dev.step(q, 2)
    ora.step(q, 2)
    assert dev.deaths_count == ora.deaths_count, "two more steps: deaths"
    parity.assert_same_state(dev, ora, what="two more steps at half the time step")
    parity.assert_same_neighbors(dev, ora, "two more steps at half the time step")


@pytest.mark.parametrize("real_bytes", [4, 8])
def test_crowded_voxel_grows_the_bucket_tables(real_bytes):
    """70 agents packed into one 30 um voxel: the fused 3D mesh update's unordered voxel
    tables (32 slots) overflow, the step finishes the dense way with larger tables, and
    the next steps use them; every step agrees with the oracle."""
    mesh, st, p = parity.rich_population(3, 700, 71, real_bytes, rates="busy")
    pos = st["positions"].reshape(-1, 3).copy()
    rng = np.random.default_rng(5)
    pos[100:170] = rng.uniform(4.0, 26.0, size=(70, 3))  # the voxel [0, 30)^3
    st["positions"] = pos.reshape(-1)
    dev, ora = parity.make_pair(mesh, st, seed=6)
    for s in (dev, ora):
        s.initialize_mesh()
    for step in range(3):
        dev.step(p, 1)
        ora.step(p, 1)
        assert dev.deaths_count == ora.deaths_count, f"step {step}: deaths"
        parity.assert_same_state(dev, ora, what=f"crowded voxel, step {step}")
        parity.assert_same_buckets(dev, ora, f"crowded voxel, step {step}")
        parity.assert_same_neighbors(dev, ora, f"crowded voxel, step {step}")


@pytest.mark.parametrize("dims", [2, 3])
def test_neighbors_with_stale_mesh(dims):
    """Neighbors rebuilt after the agents moved but before a new mesh update: the stencil is
    taken around each agent's current voxel over the old buckets, as in the reference."""
    mesh, st, p = parity.rich_population(dims, 1200, 50 + dims, 4, rates="rare")
    dev, ora = parity.make_pair(mesh, st, seed=2)
    for s in (dev, ora):
        s.initialize_mesh()
        s.update_cell_forces(p)
        s.update_positions(p)
        s.update_positions(p)
        s.update_cell_neighbors(p)
    parity.assert_same_neighbors(dev, ora, "stale mesh")


def test_fused_equals_phases():
    """The fused step and the per-phase entry points give bit-identical states."""
    mesh, st, p = parity.rich_population(3, 1500, 33, 4, rates="busy")
    a = device(mesh, st, seed=3)
    b = device(mesh, st, seed=3)
    for s in (a, b):
        s.initialize_mesh()
    for _ in range(3):
        a.step(p, 1)
        b.run_phases(p)
    parity.assert_same_state(a, b, tol=0.0, what="fused vs phases")
    parity.assert_same_neighbors(a, b, "fused vs phases")


def test_deterministic():
    """Two runs from the same state are bitwise identical (no schedule dependence)."""
    mesh, st, p = parity.rich_population(3, 2000, 44, 4, rates="busy")
    runs = []
    for _ in range(2):
        d = device(mesh, st, seed=1)
        d.initialize_mesh()
        d.step(p, 5)
        runs.append(d)
    parity.assert_same_state(runs[0], runs[1], tol=0.0, what="rerun")
    parity.assert_same_neighbors(runs[0], runs[1], "rerun")


def _flag_divisions(st, seed):
    """5% of the agents divide, 2% are flagged for removal; random unit orientations and
    polarity 0 or 1."""
    n, dims = st.count, st.dims
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    flags = np.zeros(n, np.uint8)
    flags[perm[: n // 20]] = 2  # PCM_FLAG_TO_DIVIDE
    flags[perm[n // 20: n // 20 + n // 50]] = 1  # PCM_FLAG_TO_REMOVE
    st["flags"] = flags
    o = rng.normal(size=(n, dims))
    st["orientation"] = (o / np.linalg.norm(o, axis=1, keepdims=True)).ravel()
    st["polarity"] = rng.integers(0, 2, n).astype(float)


@pytest.mark.parametrize("real_bytes", [4, 8])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_phenotype_container_update(dims, real_bytes):
    """pcm_update_cell_container_for_phenotype against the oracle: divisions (daughters at
    the box edge included), the removal loop over the old agents, the division log and
    counters; then the rest of simulate_phenotype's housekeeping (mesh, neighbors,
    simulator.cpp:208-216) and fused steps from the new population."""
    n = {1: 200, 2: 800, 3: 1500}[dims]
    mesh, st, p = parity.rich_population(dims, n, 60 + dims, real_bytes, rates="busy")
    _flag_divisions(st, 70 + dims)
    dev, ora = parity.make_pair(mesh, st, seed=4, capacity=2 * n)
    for s in (dev, ora):
        s.initialize_mesh()
        s.update_cell_container_for_phenotype(p, compute_internalized_substrates=real_bytes == 8)
    parity.assert_same_state(dev, ora, what="phenotype update")
    assert np.array_equal(dev.division_log(), ora.division_log())
    assert dev.divisions_count == ora.divisions_count == n // 20
    assert dev.deaths_count == ora.deaths_count
    for s in (dev, ora):
        s.initialize_mesh()
    parity.assert_same_buckets(dev, ora, "mesh after the phenotype update")
    parity.assert_same_neighbors(dev, ora, "neighbors after the phenotype update")
    # the steps start from the oracle's state: a daughter sits 1.5 radii from its parent,
    # where the repulsion turns a last-bit difference of the walk's cos/sin into more
    dev.upload(ora.download())
    dev.initialize_mesh()
    for step in range(2):
        dev.step(p, 1)
        ora.step(p, 1)
        parity.assert_same_state(dev, ora, what=f"step {step} after the phenotype update")
        parity.assert_same_neighbors(dev, ora, f"step {step} after the phenotype update")
        dev.upload(ora.download())
        dev.step_index = ora.step_index
        dev.initialize_mesh()


def test_phenotype_update_capacity_is_checked_first():
    """Divisions beyond the handle's capacity fail before any agent changes."""
    mesh, st, p = parity.rich_population(3, 400, 81, 4, rates="rare")
    _flag_divisions(st, 82)
    dev = device(mesh, st, seed=1, capacity=410)
    dev.initialize_mesh()
    before = dev.download()
    with pytest.raises(Exception, match="capacity"):
        dev.update_cell_container_for_phenotype(p)
    after = dev.download()
    for name, _, _ in abi.FIELDS:
        assert np.array_equal(before.arrays[name], after.arrays[name]), name


@pytest.mark.parametrize("dims", [2, 3])
def test_pack_fields_is_the_field_matrix(dims):
    """pcm_pack_fields gathers the MultiCellDS writer's device columns
    (PhysiCell_MultiCellDS.cpp:1735-1790) exactly: values as doubles, vectors padded to 3."""
    mesh, st, p = parity.rich_population(dims, 700, 90 + dims, 4, rates="busy")
    st["orientation"] = np.random.default_rng(5).normal(size=st.count * dims)
    dev = device(mesh, st, seed=2)
    dev.initialize_mesh()
    dev.step(p, 2)
    names = ["positions", "volume", "nuclear", "cytoplasmic", "fluid_fraction", "orientation", "polarity",
             "velocities", "simple_pressure", "number_of_nuclei", "damage", "total_attack_time",
             "cell_adhesion_affinities", "dead", "flags", "internalized_substrates"]
    widths = [3, 0, 0, 0, 0, 3, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0]
    got = dev.pack_fields(names, widths)
    n = dev.count
    cols = []
    for name, w in zip(names, widths):
        a = dev.get_field(name).reshape(n, -1).astype(np.float64)
        if w > a.shape[1]:
            a = np.hstack([a, np.zeros((n, w - a.shape[1]))])
        cols.append(a)
    assert np.array_equal(got, np.hstack(cols))


def test_removal_log_replays_the_compaction():
    """Replaying pcm_removal_log's positions as swap-with-last removals on the uploaded
    order gives the device's final order (ORIGIN)."""
    mesh, st, p = parity.rich_population(3, 1200, 95, 4, rates="busy")
    dev = device(mesh, st, seed=6)
    dev.initialize_mesh()
    dev.removal_log()
    dev.step(p, 3)
    order = list(range(st.count))
    for i in dev.removal_log():
        order[i] = order[-1]
        order.pop()
    assert np.array_equal(np.array(order, np.int32), dev.get_field("origin"))


def test_uploads_after_removals_read_count_rows():
    """After the container update removed agents, pcm_set_field, pcm_set_links and
    pcm_set_neighbors read exactly the current count of rows: a guard tail of invalid
    values after them is never read, and the round trip is exact."""
    mesh, st, p = parity.rich_population(3, 1200, 95, 4, rates="busy")
    dev = device(mesh, st, seed=6)
    dev.initialize_mesh()
    dev.step(p, 3)
    n = dev.count
    assert n < st.count
    for name, bad in (("maximum_number_of_attachments", 1000), ("contact_function", 200)):
        a = dev.get_field(name)
        guard = np.full(st.count, bad, dtype=a.dtype)
        guard[:n] = a
        dev._call("set_field", dev._h, abi.FIELD_ID[name], guard.ctypes.data)
        assert np.array_equal(dev.get_field(name), a)
    for which in (abi.LINKS_SPRINGS, abi.LINKS_ATTACHED):
        off, idx = dev.links(which)
        guard = np.zeros(st.count + 1, np.int32)  # decreasing offsets after the count
        guard[: n + 1] = off
        entries = np.append(idx, 0).astype(np.int32)  # referenced until the call returns
        dev._call("set_links", dev._h, which, guard.ctypes.data, entries.ctypes.data)
        off2, idx2 = dev.links(which)
        assert np.array_equal(off2, off) and np.array_equal(idx2, idx)
    off, idx = dev.neighbors()
    guard = np.zeros(st.count + 1, np.int32)
    guard[: n + 1] = off
    entries = np.append(idx, 0).astype(np.int32)
    dev._call("set_neighbors", dev._h, guard.ctypes.data, entries.ctypes.data)
    off2, idx2 = dev.neighbors()
    assert np.array_equal(off2, off) and np.array_equal(idx2, idx)
