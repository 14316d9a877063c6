"""The benchmark workloads at their full sizes, device against the oracle (single-thread,
keyed random streams) from identical states: the 900-agent 2D disc, the 14.3k-agent 3D
ball (FP32 and FP64) and the 88.7k-agent 3D ball with springs and interactions.  A few
fused steps each; reals within 1e-5 (FP32) / 1e-12 (FP64) of max(|value|, 1), integer
fields, flags, link lists and neighbor lists exactly.  Size-independent invariants of the
mesh are checked on the device state as well."""
from __future__ import annotations

import numpy as np
import pytest

import parity
from paraphysicell_amd import abi, synth

pytestmark = pytest.mark.gpu


def _check_mesh_invariants(dev, st_mesh):
    """Every agent sits in exactly one bucket, the bucket of its voxel, buckets descend."""
    counts, agents = dev.buckets()
    n = dev.count
    assert counts.sum() == n and np.array_equal(np.sort(agents), np.arange(n))
    pos = dev.get_field("positions").reshape(n, st_mesh.dims).astype(np.float64)
    g = st_mesh.grid_shape()
    start = np.concatenate([[0], np.cumsum(counts)])
    for v in np.nonzero(counts)[0][:: max(1, len(np.nonzero(counts)[0]) // 300)]:
        members = agents[start[v]:start[v + 1]]
        assert np.all(np.diff(members) < 0)
        c = [v % g[0], (v // g[0]) % g[1], v // (g[0] * g[1])]
        for k in range(st_mesh.dims):
            lo = st_mesh.bbox_min[k] + c[k] * st_mesh.voxel[k]
            inside = (pos[members, k] >= lo - 1e-3) & (pos[members, k] <= lo + st_mesh.voxel[k] + 1e-3)
            edge = (c[k] == 0) | (c[k] == g[k] - 1)
            assert np.all(inside | edge)


@pytest.mark.parametrize("case", ["disc_900_2d", "disc_12k_2d", "ball_14k_3d_f32", "ball_14k_3d_f64", "ball_88k_3d",
                                  "sparse_ball_9k_3d_f32", "sparse_ball_9k_3d_f64"])
def test_workload_full_size(case):
    print(f"full-size workload {case}", flush=True)
    if case == "disc_900_2d":
        mesh, st, params = synth.config0(4)
        steps = 10
    elif case == "disc_12k_2d":
        # the 900-agent disc's density at 12,000 agents: above the size where 2D neighbor
        # lists are built one lane per agent
        mesh, st, params = synth.ball_population(2, 12000, 400.0 * (12000 / 900) ** 0.5, 2000, 9, 4)
        steps = 5
    elif case == "ball_14k_3d_f32":
        mesh, st, params = synth.config1(4)
        steps = 3
    elif case == "ball_14k_3d_f64":
        mesh, st, params = synth.config1(8)
        steps = 3
    elif case.startswith("sparse_ball_9k_3d"):
        # the ensemble's 3D density (about two agents per occupied voxel): the neighbor search
        # stages short stencils with the sparse-population tile
        mesh, st, params = synth.config3_member(1, 9000, 4 if case.endswith("f32") else 8)
        steps = 4
    else:
        mesh, st, params = synth.config2(4)
        steps = 2
    dev, ora = parity.make_pair(mesh, st, seed=5)
    for s in (dev, ora):
        s.initialize_mesh()
    parity.assert_same_neighbors(dev, ora, f"{case} initial")
    for s in (dev, ora):
        s.step(params, steps)
    if case.startswith("sparse"):
        occupied = dev.stats()["occupied_voxels"]
        assert dev.count < 4 * occupied, "not a sparse population"
    parity.assert_same_state(dev, ora, what=case)
    parity.assert_same_neighbors(dev, ora, case)
    parity.assert_same_buckets(dev, ora, case)
    _check_mesh_invariants(dev, mesh)
