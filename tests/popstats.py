"""Population-level statistics of a mechanics run, and their comparison across seeds.

Long trajectories of the device (counter-based Philox streams) and of the reference
(thread-local mt19937, restated by the oracle's "mt" mode) differ draw by draw, so runs
are compared by the statistics a model user reads off a population, per seed, over
several seeds started from identical populations:

- shape: radius of gyration, quantiles of the distance from the population's start centre
  (the radial density profile), mean squared displacement since the start (motility);
- contacts: mean neighbor count, the fraction without neighbors and neighbor-count
  quantiles;
- pressure: mean and 95th percentile of simple_pressure;
- springs: mean springs per agent, the fraction without springs, spring-degree quantiles;
- interactions: mean damage, the fraction of damaged agents, agents removed.

``compare`` pairs the runs of each seed (same initial population on both sides) and
tests every statistic's mean paired difference against its spread over the seeds.
"""
from __future__ import annotations

import numpy as np

QUANTILES = (0.1, 0.25, 0.5, 0.75, 0.9, 0.99)


def population_statistics(sim, initial_positions: np.ndarray, dims: int, n0: int, springs: bool) -> dict:
    """Statistics of `sim`'s current state; `initial_positions` (n0 x dims) are the
    positions at upload, indexed by each agent's ORIGIN field."""
    n = sim.count
    pos = sim.get_field("positions").reshape(n, dims).astype(np.float64)
    origin = sim.get_field("origin")
    start = initial_positions.astype(np.float64)
    centre = start.mean(axis=0)
    out = {}
    centroid = pos.mean(axis=0)
    out["radius_of_gyration"] = float(np.sqrt(np.mean(np.sum((pos - centroid) ** 2, axis=1))))
    r = np.sqrt(np.sum((pos - centre) ** 2, axis=1))
    for q in QUANTILES:
        out[f"radial_q{int(q * 100)}"] = float(np.quantile(r, q))
    disp = pos - start[origin]
    out["msd"] = float(np.mean(np.sum(disp ** 2, axis=1)))
    off, _ = sim.neighbors()
    deg = np.diff(off)
    out["neighbors_mean"] = float(deg.mean())
    out["neighbors_none"] = float(np.mean(deg == 0))
    for q in (0.1, 0.5, 0.9):
        out[f"neighbors_q{int(q * 100)}"] = float(np.quantile(deg, q))
    p = sim.get_field("simple_pressure").astype(np.float64)
    out["pressure_mean"] = float(p.mean())
    out["pressure_p95"] = float(np.quantile(p, 0.95))
    if springs:
        soff, _ = sim.links(0)
        sd = np.diff(soff)
        out["springs_mean"] = float(sd.mean())
        out["springs_none"] = float(np.mean(sd == 0))
        for q in (0.5, 0.9):
            out[f"springs_q{int(q * 100)}"] = float(np.quantile(sd, q))
        dmg = sim.get_field("damage").astype(np.float64)
        out["damage_mean"] = float(dmg.mean())
        out["damaged_fraction"] = float(np.mean(dmg > 0))
        out["removed"] = float(n0 - n)
    return out


# |t| of a paired difference over 8 seeds (7 degrees of freedom) above this fails: the
# two-sided tail beyond 7.0 is 2e-4, so ~30 statistics per workload give a false alarm in
# under 1% of runs, while a statistic shifted by about 2.5 standard errors of the
# per-seed difference, times sqrt(8), is caught
T_LIMIT = 7.0


def compare(dev: list[dict], ref: list[dict], t_limit: float = T_LIMIT, rel_floor: float = 1e-6):
    """Paired comparison per statistic: mean of (dev - ref) over seeds against its standard
    error.  A statistic whose paired differences are all within rel_floor of the reference
    scale passes outright (identical values on both sides, e.g. integer quantiles).
    Returns a list of (name, mean difference, t, reference mean) for failing statistics."""
    assert len(dev) == len(ref) and len(dev) >= 3
    bad = []
    for name in ref[0]:
        a = np.array([d[name] for d in dev])
        b = np.array([r[name] for r in ref])
        d = a - b
        scale = max(float(np.mean(np.abs(b))), 1e-12)
        if np.all(np.abs(d) <= rel_floor * scale):
            continue
        se = float(np.std(d, ddof=1) / np.sqrt(len(d)))
        t = float(np.mean(d)) / se if se > 0 else float("inf")
        if abs(t) > t_limit:
            bad.append((name, float(np.mean(d)), t, float(np.mean(b))))
    return bad
