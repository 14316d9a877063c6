"""Synthetic agent populations of the benchmark configurations.

Positions are uniform in a disc (2D) or ball (3D) by inverse-CDF radius and uniform
direction; phenotypes start from the reference defaults (src/data_storage_phenotype.cpp)
with the coefficients each configuration names.  All draws come from numpy's PCG64 with
the configuration seed, so the CPU and GPU legs start from the same state.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .state import AgentState

DEFAULT_RADIUS = 8.412710547954228


def uniform_ball(rng, n: int, dims: int, radius: float) -> np.ndarray:
    direction = rng.normal(size=(n, dims))
    direction /= np.linalg.norm(direction, axis=1, keepdims=True)
    r = radius * rng.uniform(0.0, 1.0, n) ** (1.0 / dims)
    return direction * r[:, None]


def ball_population(dims: int, n: int, radius: float, box: int, seed: int, real_bytes: int = 4, defs: int = 1,
                    subs: int = 1, radius_sd: float = 0.0, motile: bool = True, wall: bool = True,
                    springs: bool = False, dt: float = 0.1):
    """n default agents uniform in a ball of `radius` inside the box [-box, box]^dims."""
    rng = np.random.default_rng(seed)
    mesh = abi.MeshSpec(dims, (-box,) * 3, (box,) * 3)
    st = AgentState(mesh, real_bytes, defs, subs, n).defaults()
    st["positions"] = uniform_ball(rng, n, dims, radius).reshape(-1)
    if radius_sd > 0:
        st["radius"] = np.clip(rng.normal(8.41, radius_sd, n), 6.0, 11.0)
    st["is_motile"] = 1 if motile else 0
    params = abi.Params(dt=dt, virtual_wall=wall, springs=springs)
    return mesh, st, params


def config0(real_bytes: int = 4, seed: int = 0):
    """2D, 900 agents in a disc of 400 um, box [-1000, 1000]^2, motile, virtual wall on."""
    return ball_population(2, 900, 400.0, 1000, seed, real_bytes, motile=True, wall=True)


def config1(real_bytes: int = 4, seed: int = 1, n: int = 14300):
    """3D, 14300 agents in a ball of 200 um, radii N(8.41, 0.5) clipped to [6, 11], BM on."""
    return ball_population(3, n, 200.0 * (n / 14300) ** (1 / 3), 1000, seed, real_bytes, radius_sd=0.5,
                           motile=True, wall=True)


def config2(real_bytes: int = 4, seed: int = 2, n: int = 88700, ball: float = 380.0, box: int = 2000):
    """3D, 88700 agents in a ball of 380 um, two types 70/30 with springs and interactions.

    Springs: elastic 0.05, attachment 0.01/min, detachment 0.001/min, at most 12.
    Type B (index 1) attacks type A (index 0) at 0.005/min with damage rate 1 and eats dead
    cells at 0.01/min; 1% of the agents start dead (not motile) so phagocytosis occurs.
    """
    scale = (n / 88700) ** (1 / 3)
    mesh, st, params = ball_population(3, n, ball * scale, box, seed, real_bytes, defs=2, subs=1, motile=True,
                                       wall=True, springs=True)
    rng = np.random.default_rng(seed + 1000)
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: int(round(0.3 * n))]] = 1
    st["cell_definition_index"] = types
    st["attachment_elastic_constant"] = 0.05
    st["attachment_rate"] = 0.01
    st["detachment_rate"] = 0.001
    st["maximum_number_of_attachments"] = 12
    attack = st.view("attack_rates")
    attack[types == 1, 0] = 0.005
    st["damage_rate"] = 1.0
    st["dead_phagocytosis_rate"] = np.where(types == 1, 0.01, 0.0)
    dead = np.zeros(n, np.uint8)
    dead[rng.permutation(n)[: int(round(0.01 * n))]] = 1
    st["dead"] = dead
    st["is_motile"] = np.where(dead == 1, 0, 1)
    return mesh, st, params


def config3_sizes(count: int = 64, lo: int = 500, hi: int = 20000, seed: int = 0):
    """Per-replicate agent counts, log-uniform in [lo, hi] (seeded)."""
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(count)]


def config3_member(s: int, n: int, real_bytes: int = 4):
    """Replicate s of the 64-member ensemble: 2D if s is even, else 3D, at the density of the
    900-agent 2D workload (disc of 400 um; the 3D density is its 3/2 power), in a box of
    2.5 radii rounded up to 20 um, with that workload's coefficients (motile, wall on)."""
    dims = 2 if s % 2 == 0 else 3
    rho2 = 900 / (math.pi * 400.0 ** 2)
    if dims == 2:
        radius = 400.0 * math.sqrt(n / 900)
    else:
        rho3 = rho2 ** 1.5
        radius = (3 * n / (4 * math.pi * rho3)) ** (1 / 3)
    box = int(math.ceil(2.5 * radius / 20.0) * 20)
    return ball_population(dims, n, radius, box, s, real_bytes, motile=True, wall=True)


def serving_request(k: int, seed: int = 0):
    """Request k of the served workload: log-uniform 1k-200k agents, 2D or 3D at equal odds,
    with the 88.7k-agent workload's physics (two types, springs, attack, dead phagocytosis).
    3D populations keep that workload's density (ball of 380 um scaled by (N/88700)^(1/3));
    2D ones use the 900-agent 2D workload's density (disc of 400 um scaled by
    (N/900)^(1/2)), since a disc at the 3D count per area would overlap every cell."""
    rng = np.random.default_rng([seed, k])
    n = int(round(math.exp(rng.uniform(math.log(1000), math.log(200000)))))
    dims = 2 if rng.uniform() < 0.5 else 3
    if dims == 3:
        ball = 380.0 * (n / 88700) ** (1 / 3)
    else:
        ball = 400.0 * math.sqrt(n / 900)
    box = int(math.ceil(2.5 * ball / 20.0) * 20)
    mesh, st, params = ball_population(dims, n, ball, box, int(rng.integers(1 << 30)), 4, defs=2, subs=1,
                                       motile=True, wall=True, springs=True)
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: int(round(0.3 * n))]] = 1
    st["cell_definition_index"] = types
    st["attachment_elastic_constant"] = 0.05
    st["attachment_rate"] = 0.01
    st["detachment_rate"] = 0.001
    st["maximum_number_of_attachments"] = 12
    st.view("attack_rates")[types == 1, 0] = 0.005
    st["damage_rate"] = 1.0
    st["dead_phagocytosis_rate"] = np.where(types == 1, 0.01, 0.0)
    dead = np.zeros(n, np.uint8)
    dead[rng.permutation(n)[: int(round(0.01 * n))]] = 1
    st["dead"] = dead
    st["is_motile"] = np.where(dead == 1, 0, 1)
    return mesh, st, params
