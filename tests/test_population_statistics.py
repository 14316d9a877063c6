"""Population-level parity with the reference's own random stream.

The fixture tests/golden/population_stats.npz holds the statistics of tests/popstats.py
for the oracle in its "mt" mode (the reference's thread-local mt19937 draws): the 900-agent
2D disc after 600 steps and the 88.7k-agent 3D ball after 100 steps, 8 seeds each.  The
device, whose Philox streams differ draw by draw, must match every statistic by the paired
test of popstats.compare (|t| <= 7 over 8 seeds).

CPU checks pin the fixture and the test itself: the oracle reproduces the fixture's first
disc seed exactly; the oracle in the device's keyed-Philox mode passes the same comparison
on the disc (the two streams agree in distribution); and the comparison fails for a disc
whose motility is wrong in the ways a plausible mistake would make it (persistence time
doubled, i.e. half the redraw probability; speed off by 15%).
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from golden.make_population_stats import oracle_factory, run_stats, workloads
from popstats import compare

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "population_stats.npz")


def _fixture(name):
    z = np.load(FIXTURE)
    names = [str(x) for x in z[f"{name}__names"]]
    rows = [dict(zip(names, map(float, v))) for v in z[f"{name}__values"]]
    meta = json.loads(str(z["meta"]))
    return rows, meta


def test_fixture_reproduces_on_the_oracle():
    rows, meta = _fixture("disc_900_2d")
    make, steps, springs, threads = workloads()["disc_900_2d"]
    assert threads == 1 and meta["disc_900_2d"]["steps"] == steps
    got = run_stats(oracle_factory(1), make, meta["seeds"][0], steps, springs)
    assert got == rows[0]


def test_keyed_stream_matches_the_reference_stream_on_the_oracle():
    rows, meta = _fixture("disc_900_2d")
    make, steps, springs, _ = workloads()["disc_900_2d"]
    keyed = [run_stats(oracle_factory(1, rng="keyed"), make, s, steps, springs) for s in meta["seeds"]]
    bad = compare(keyed, rows)
    assert not bad, bad


@pytest.mark.parametrize("mistake", ["persistence_doubled", "speed_plus_15_percent"])
def test_comparison_catches_a_motility_mistake(mistake):
    rows, meta = _fixture("disc_900_2d")
    make, steps, springs, _ = workloads()["disc_900_2d"]

    def wrong(s):
        mesh, st, params = make(s)
        if mistake == "persistence_doubled":
            st["persistence_time"] = 2.0 * st["persistence_time"]
        else:
            st["migration_speed"] = 1.15 * st["migration_speed"]
        return mesh, st, params

    got = [run_stats(oracle_factory(1), wrong, s, steps, springs) for s in meta["seeds"]]
    bad = compare(got, rows)
    assert any(name == "msd" for name, *_ in bad), bad


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["disc_900_2d", "ball_88k_3d"])
def test_device_matches_the_reference_stream(name):
    from paraphysicell_amd.mechanics import DeviceSim

    rows, meta = _fixture(name)
    make, steps, springs, _ = workloads()[name]

    def device(mesh, st, seed):
        return DeviceSim(mesh, st.real_bytes, st.defs, st.subs, capacity=st.count, seed=seed)

    print(f"population statistics: {name}, {len(meta['seeds'])} seeds, {steps} steps", flush=True)
    got = [run_stats(device, make, s, steps, springs) for s in meta["seeds"]]
    bad = compare(got, rows)
    assert not bad, bad
