"""Comparison helpers: device library vs oracle on identical initial states."""
from __future__ import annotations

import numpy as np

from oracle.pyoracle import OracleSim
from paraphysicell_amd import abi
from paraphysicell_amd.mechanics import DeviceSim

TOL = {4: 1e-5, 8: 1e-12}  # |gpu - oracle| <= tol * max(|oracle|, 1)


def make_pair(mesh, st, seed=7, capacity=None):
    cap = capacity or max(st.count, 1)
    dev = DeviceSim(mesh, st.real_bytes, st.defs, st.subs, capacity=cap, seed=seed)
    ora = OracleSim(mesh, st.real_bytes, st.defs, st.subs, capacity=cap, seed=seed)
    dev.upload(st)
    ora.upload(st)
    return dev, ora


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def assert_same_state(dev, ora, tol=None, fields=None, what=""):
    """Every field: bit-exact for integer/byte fields, within tol for reals; lists exact."""
    n_d, n_o = dev.count, ora.count
    assert n_d == n_o, f"{what}: agent count {n_d} != {n_o}"
    tol = TOL[dev.real_bytes] if tol is None else tol
    for name, ftype, _ in abi.FIELDS:
        if fields is not None and name not in fields:
            continue
        a, b = dev.get_field(name), ora.get_field(name)
        if ftype == abi.T_REAL:
            e = rel_err(a, b)
            assert e <= tol, f"{what}: field {name} differs, max rel err {e:.3e} > {tol}"
        else:
            assert np.array_equal(a, b), f"{what}: field {name} differs at {np.flatnonzero(a != b)[:10]}"
    for which in (abi.LINKS_SPRINGS, abi.LINKS_ATTACHED):
        (od, idd), (oo, ido) = dev.links(which), ora.links(which)
        assert np.array_equal(od, oo) and np.array_equal(idd, ido), f"{what}: link lists {which} differ"


def assert_same_neighbors(dev, ora, what=""):
    od, idd = dev.neighbors()
    oo, ido = ora.neighbors()
    assert np.array_equal(od, oo), f"{what}: neighbor counts differ"
    assert np.array_equal(idd, ido), f"{what}: neighbor lists differ"


def assert_same_buckets(dev, ora, what=""):
    cd, ad = dev.buckets()
    co, ao = ora.buckets()
    assert np.array_equal(cd, co), f"{what}: bucket sizes differ"
    assert np.array_equal(ad, ao), f"{what}: bucket contents differ"


def rich_population(dims, n, seed, real_bytes, box=150, fill=0.55, spring_cap=3, rates="busy"):
    """A small dense population exercising every branch of the mechanics step.

    Two cell types; 10% non-movable; 5% dead; pre-existing springs (some duplicated);
    attached cells with contact functions on half of the agents; chemotaxis models on a
    random gradient field;
    attachment/detachment, attack, phagocytosis and fusion rates high enough that
    capacity limits and interaction conflicts occur; a few agents drift out of the box.
    """
    rng = np.random.default_rng(seed)
    mesh = abi.MeshSpec(dims, (-box,) * 3, (box,) * 3)
    from paraphysicell_amd.state import AgentState

    st = AgentState(mesh, real_bytes, 2, 2, n).defaults()
    pos = rng.uniform(-box * fill, box * fill, size=(n, dims))
    # a few agents right at the wall, moving out
    k = max(1, n // 100)
    pos[:k, 0] = box - 0.5
    st["positions"] = pos.reshape(-1)
    vel = rng.normal(0, 0.5, (n, dims))
    vel[:k, 0] = 50.0
    st["velocities"] = vel.reshape(-1)
    st["previous_velocities"] = rng.normal(0, 0.5, n * dims)
    st["radius"] = np.clip(rng.normal(8.41, 0.8, n), 6, 11)
    st["cell_definition_index"] = rng.integers(0, 2, n)
    st["is_movable"] = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    st["is_motile"] = (rng.uniform(size=n) > 0.2).astype(np.uint8)
    st["migration_bias"] = rng.uniform(0, 1, n) * (rng.uniform(size=n) < 0.3)
    bias_dir = rng.normal(size=(n, dims))
    st["migration_bias_direction"] = (bias_dir / np.linalg.norm(bias_dir, axis=1, keepdims=True)).reshape(-1)
    st["motility_vector"] = rng.normal(0, 0.5, n * dims)
    st["persistence_time"] = rng.uniform(0.05, 2.0, n)
    st["migration_speed"] = rng.uniform(0.5, 2.0, n)
    st["cell_adhesion_affinities"] = rng.uniform(0.2, 1.5, n * 2)
    st["cell_BM_repulsion_strength"] = 100.0
    st["maximum_number_of_attachments"] = spring_cap
    st["attachment_elastic_constant"] = rng.uniform(0.005, 0.05, n)
    busy = rates == "busy"
    st["attachment_rate"] = rng.uniform(0, 4.0 if busy else 0.05, n)
    st["detachment_rate"] = rng.uniform(0, 1.0 if busy else 0.01, n)
    st["dead"] = (rng.uniform(size=n) < 0.05).astype(np.uint8)
    st["dead_phagocytosis_rate"] = rng.uniform(0, 3.0 if busy else 0.02, n)
    st["live_phagocytosis_rates"] = rng.uniform(0, 0.3 if busy else 0.002, n * 2)
    st["attack_rates"] = rng.uniform(0, 2.0 if busy else 0.02, n * 2)
    st["immunogenicities"] = rng.uniform(0, 1, n * 2)
    st["fusion_rates"] = rng.uniform(0, 0.2 if busy else 0.002, n * 2)
    st["damage_rate"] = rng.uniform(0.5, 1.5, n)
    st["internalized_substrates"] = rng.uniform(0, 1, n * 2)
    st["fraction_transferred_when_ingested"] = rng.uniform(0, 1, n * 2)
    # symmetric initial springs between nearby agents, a few duplicated
    lists = [[] for _ in range(n)]
    order = rng.permutation(n)
    for a, b in zip(order[0::2], order[1::2]):
        if len(lists[a]) < spring_cap and len(lists[b]) < spring_cap:
            lists[a].append(int(b))
            lists[b].append(int(a))
            if rng.uniform() < 0.1 and len(lists[a]) < spring_cap and len(lists[b]) < spring_cap:
                lists[a].append(int(b))
                lists[b].append(int(a))
    st.set_links(abi.LINKS_SPRINGS, lists)
    att = [[] for _ in range(n)]
    for a, b in zip(order[1::3], order[2::3]):
        att[a].append(int(b))
        att[b].append(int(a))
    # a few hubs attached to more cells than a lane group holds (lists stay symmetric, as
    # cell::attach_cells makes them)
    for h in order[: max(1, n // 100)]:
        for x in rng.choice(n, size=min(n, 13), replace=False):
            if x != h:
                att[h].append(int(x))
                att[int(x)].append(int(h))
    st.set_links(abi.LINKS_ATTACHED, att)
    # half of the agents carry the standard elastic contact function
    st["contact_function"] = (rng.uniform(size=n) < 0.5).astype(np.uint8)
    # chemotaxis: every direction update model over a random gradient field on an
    # anisotropic diffusion grid (a few zero gradients, so normalisation meets length 0)
    st["migration_bias_direction_function"] = rng.integers(0, 4, n).astype(np.uint8)
    st["chemotaxis_index"] = rng.integers(0, 2, n)
    st["chemotaxis_direction"] = rng.choice([-1, 1], n)
    st["chemotactic_sensitivities"] = rng.uniform(-1, 1, n * 2)
    dvox = (25, 20, 15)
    voxels = int(np.prod(mesh.with_voxel(dvox).grid_shape()))
    grad = rng.normal(0, 1, (voxels, 2, dims))
    grad[rng.uniform(size=voxels) < 0.05] = 0.0
    st.gradients = (dvox, grad.astype(st.real).ravel())
    params = abi.Params(dt=0.1, virtual_wall=True, springs=True)
    return mesh, st, params
