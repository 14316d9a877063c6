"""Writes tests/golden/population_stats.npz: population statistics (tests/popstats.py) of the
oracle in its "mt" random mode, the reference's own stream (a thread-local mt19937 per
OpenMP thread with uniform_real_distribution draws, src/random.cpp:8-49), on two of the
benchmark workloads over 8 seeds each:

- disc_900_2d: the 900-agent 2D disc (motility, membrane), 600 mechanics steps, 1 thread;
- ball_88k_3d: the 88.7k-agent 3D ball (springs, attack, dead phagocytosis), 100 steps,
  8 OpenMP threads (the reference's threaded draw order).

Seed s uses the synthetic population of seed s and random seed 1000 + s on both the
oracle and the device.  Run in the build container (the reference itself is never run):

    python tests/golden/make_population_stats.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle.pyoracle import OracleSim  # noqa: E402
from paraphysicell_amd import synth  # noqa: E402
from popstats import population_statistics  # noqa: E402

SEEDS = list(range(8))
OUT = os.path.join(HERE, "population_stats.npz")


def workloads():
    """name -> (population maker, steps, has springs, oracle threads)"""
    return {
        "disc_900_2d": (lambda s: synth.config0(4, seed=s), 600, False, 1),
        "ball_88k_3d": (lambda s: synth.config2(4, seed=s), 100, True, 8),
    }


def run_stats(sim_factory, make, seed, steps, springs):
    mesh, st, params = make(seed)
    sim = sim_factory(mesh, st, 1000 + seed)
    sim.upload(st)
    sim.initialize_mesh()
    sim.step(params, steps)
    start = st["positions"].reshape(st.count, mesh.dims)
    out = population_statistics(sim, start, mesh.dims, st.count, springs)
    sim.close()
    return out


def oracle_factory(threads, rng="mt"):
    def make(mesh, st, seed):
        return OracleSim(mesh, st.real_bytes, st.defs, st.subs, capacity=st.count, seed=seed, threads=threads, rng=rng)

    return make


def main():
    arrays = {}
    meta = {"seeds": SEEDS, "random_seed": "1000 + seed", "rng": "mt"}
    for name, (make, steps, springs, threads) in workloads().items():
        rows = []
        t0 = time.time()
        for s in SEEDS:
            rows.append(run_stats(oracle_factory(threads), make, s, steps, springs))
            print(f"{name} seed {s}: {time.time() - t0:.1f} s", flush=True)
        names = list(rows[0])
        arrays[f"{name}__names"] = np.array(names)
        arrays[f"{name}__values"] = np.array([[r[k] for k in names] for r in rows])
        meta[name] = {"steps": steps, "threads": threads, "springs": springs}
    arrays["meta"] = np.array(json.dumps(meta))
    np.savez(OUT, **arrays)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
