"""Initial states of the reference's unit tests (test/test.cpp) and of small random cases.

Every builder returns (MeshSpec, AgentState, Params).  The reference fixtures use a
diffusion mesh with 20 um voxels and the default 30 um mechanics voxel
(test/test.cpp:13-50, src/environment.h:10); only the bounding box and dims matter here.
"""
from __future__ import annotations

import numpy as np

from paraphysicell_amd import abi
from paraphysicell_amd.state import AgentState


def zero_cell_state(mesh, real_bytes, defs, subs, n):
    """cont.create(): zero phenotype, movable, state defaults (src/cell.cpp:33-60)."""
    st = AgentState(mesh, real_bytes, defs, subs, n)
    st["is_movable"] = 1
    st["number_of_nuclei"] = 1
    return st


def velocity_simple(dims, real_bytes=4):
    """host_velocity_solver.simple (test/test.cpp:117-191)."""
    mesh = abi.MeshSpec(dims, (0, 0, 0), (80, 80, 80))
    st = zero_cell_state(mesh, real_bytes, 3, 2, 3)
    for c, (offset, cdef) in enumerate([(0, 0), (10, 1)]):
        st.view("positions")[c, :] = [1 + offset + i for i in range(dims)]
        st["cell_definition_index"][c] = cdef
        st["radius"][c] = 4 + offset
        st["cell_cell_adhesion_strength"][c] = 5 + offset
        st["cell_cell_repulsion_strength"][c] = 6 + offset
        st["relative_maximum_adhesion_distance"][c] = 7 + offset
        st.view("cell_adhesion_affinities")[c, :] = [8 + offset, 9 + offset, 10 + offset]
    st.view("positions")[2, :] = 70
    return mesh, st, abi.Params()


VELOCITY_SIMPLE_EXPECTED = {3: 55.61362893236258, 2: 69.556601146394243, 1: 100.15967659502964}


def velocity_complex(dims, real_bytes=4):
    """host_velocity_solver.complex (test/test.cpp:265-320)."""
    mesh, st, p = velocity_simple(dims, real_bytes)
    r = 2 if dims == 1 else 4
    for c, pos in enumerate([1, 4, 7]):
        st["radius"][c] = r
        st["relative_maximum_adhesion_distance"][c] = 1
        st.view("positions")[c, :] = pos
    return mesh, st, p


def brute_force_velocities(st):
    """compute_expected_velocities (test/test.cpp:193-263): all pairs, no cut-offs, FP64."""
    n, d = st.count, st.dims
    pos = st.view("positions").astype(np.float64)
    r = st["radius"].astype(np.float64)
    R = st["cell_cell_repulsion_strength"].astype(np.float64)
    A = st["cell_cell_adhesion_strength"].astype(np.float64)
    rel = st["relative_maximum_adhesion_distance"].astype(np.float64)
    aff = st.view("cell_adhesion_affinities").astype(np.float64)
    t = st["cell_definition_index"]
    out = np.zeros((n, d))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            diff = pos[i] - pos[j]
            dist = np.sqrt((diff * diff).sum())
            rep = (1 - dist / (r[i] + r[j])) ** 2 * np.sqrt(R[i] * R[j])
            adh = (1 - dist / (rel[i] * r[i] + rel[j] * r[j])) ** 2
            adh *= np.sqrt(A[i] * A[j] * aff[i, t[j]] * aff[j, t[i]])
            out[i] += (rep - adh) / dist * diff
    return out


def coords_from_index(index, grid, dims):
    """get_coords_from_index (test/test.cpp:322-331)."""
    if dims == 1:
        return (index, 0, 0)
    if dims == 2:
        return (index % grid[0], index // grid[1], 0)
    return (index % grid[0], (index // grid[0]) % grid[1], index // (grid[0] * grid[1]))


def neighbors_grid(dims, per_voxel, adhesion, real_bytes=4):
    """host_neighbors_solver.grid (test/test.cpp:333-400)."""
    mesh = abi.MeshSpec(dims, (0, 0, 0), (200, 200, 200))
    grid = mesh.grid_shape()
    voxels = grid[0] * grid[1] * grid[2]
    st = zero_cell_state(mesh, real_bytes, 3, 2, voxels * per_voxel)
    pos = st.view("positions")
    k = 0
    for v in range(voxels):
        c = coords_from_index(v, grid, dims)
        centre = [mesh.bbox_min[a] + mesh.voxel[a] * c[a] + mesh.voxel[a] / 2 for a in range(3)]
        for _ in range(per_voxel):
            pos[k, :] = centre[:dims]
            k += 1
    st["radius"] = 9
    st["relative_maximum_adhesion_distance"] = adhesion
    return mesh, st, abi.Params()


def expected_grid_neighbor_count(coord, grid, per_voxel, adhesion):
    """Closed form of test/test.cpp:382-399."""
    expected = per_voxel - 1
    if adhesion == 3:
        for z in (-1, 0, 1):
            for y in (-1, 0, 1):
                for x in (-1, 0, 1):
                    c = (coord[0] + x, coord[1] + y, coord[2] + z)
                    if (x, y, z) == (0, 0, 0):
                        continue
                    if all(0 <= c[a] < grid[a] for a in range(3)):
                        expected += per_voxel
    return expected


def interactions_case(kind, real_bytes=4):
    """host_interactions_solver.{attack,dead_phagocytosis,phagocytosis,fusion} (test/test.cpp:402-560)."""
    mesh = abi.MeshSpec(3, (0, 0, 0), (100, 100, 100))
    st = AgentState(mesh, real_bytes, 3, 2, 3).defaults()
    st["positions"] = 10
    if kind == "attack":
        st.view("attack_rates")[0, 0] = 1e9
        st.view("attack_rates")[1, 0] = 0
        st.view("immunogenicities")[1, 0] = 0
        st.view("attack_rates")[2, 0] = 0
        st.view("immunogenicities")[2, 0] = 1
    elif kind == "dead_phagocytosis":
        st["dead_phagocytosis_rate"][0] = 1e9
        st["dead"][1] = 1
        st["dead"][2] = 0
    elif kind == "phagocytosis":
        st["cell_definition_index"][:] = [0, 1, 0]
        st.view("live_phagocytosis_rates")[0, :2] = [0, 1e9]
    elif kind == "fusion":
        st["cell_definition_index"][:] = [0, 1, 0]
        st.view("fusion_rates")[0, :2] = [0, 1e9]
    return mesh, st, abi.Params()


# ------------------------------------------------------------------ random populations
def random_population(dims, n, seed, real_bytes=4, defs=1, subs=1, box=200, fill=0.6, motile=True,
                      radius_jitter=0.5):
    """Agents uniform in a cube filling `fill` of a [-box, box]^dims domain."""
    rng = np.random.default_rng(seed)
    mesh = abi.MeshSpec(dims, (-box,) * 3, (box,) * 3)
    st = AgentState(mesh, real_bytes, defs, subs, n).defaults()
    st["positions"] = rng.uniform(-box * fill, box * fill, size=n * dims)
    st["radius"] = np.clip(rng.normal(8.41, radius_jitter, n), 6, 11)
    st["is_motile"] = 1 if motile else 0
    st["cell_definition_index"] = rng.integers(0, defs, n)
    st["velocities"] = rng.normal(0, 0.5, n * dims)
    st["previous_velocities"] = rng.normal(0, 0.5, n * dims)
    st["motility_vector"] = rng.normal(0, 0.5, n * dims)
    return mesh, st
