"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): small
populations with every branch of the step active, after 1 and after 10 fused steps.

CPU: the oracle still reproduces them bit for bit (guards the checker itself).
GPU: the device library reproduces them through the C-ABI, reals within 1e-5 (FP32) /
1e-12 (FP64) of max(|value|, 1), integers, flags and lists exactly.
"""
from __future__ import annotations

import glob
import json
import os

import numpy as np
import pytest

from paraphysicell_amd import abi
from paraphysicell_amd.state import AgentState

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "rich_*.npz")))
TOL = {4: 1e-5, 8: 1e-12}


def load_case(path):
    z = np.load(path, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    mesh = abi.MeshSpec(meta["dims"], tuple(meta["bbox_min"]), tuple(meta["bbox_max"]), tuple(meta["voxel"]))
    st = AgentState(mesh, meta["real_bytes"], meta["defs"], meta["subs"], meta["count"])
    for name, _, _ in abi.FIELDS:
        st.arrays[name] = z[f"in_field_{name}"].copy()
    st.links[abi.LINKS_SPRINGS] = (z["in_springs_off"], z["in_springs_idx"])
    st.links[abi.LINKS_ATTACHED] = (z["in_attached_off"], z["in_attached_idx"])
    if "in_gradients" in z:
        st.gradients = (tuple(int(v) for v in z["in_gradient_voxel"]), z["in_gradients"])
    params = abi.Params(dt=meta["dt"], virtual_wall=meta["virtual_wall"], springs=meta["springs"])
    return z, meta, mesh, st, params


def check(sim, z, prefix, real_bytes, exact):
    assert sim.count == int(z[f"{prefix}count"][0])
    assert sim.deaths_count == int(z[f"{prefix}deaths"][0])
    for name, ftype, _ in abi.FIELDS:
        got, want = sim.get_field(name), z[f"{prefix}field_{name}"]
        if ftype == abi.T_REAL and not exact:
            err = np.max(np.abs(got.astype(np.float64) - want) / np.maximum(np.abs(want), 1.0), initial=0.0)
            assert err <= TOL[real_bytes], f"{prefix}{name}: {err:.3e}"
        else:
            assert np.array_equal(got, want), f"{prefix}{name}"
    for which, label in ((abi.LINKS_SPRINGS, "springs"), (abi.LINKS_ATTACHED, "attached")):
        off, idx = sim.links(which)
        assert np.array_equal(off, z[f"{prefix}{label}_off"]) and np.array_equal(idx, z[f"{prefix}{label}_idx"])
    off, idx = sim.neighbors()
    assert np.array_equal(off, z[f"{prefix}neighbors_off"]) and np.array_equal(idx, z[f"{prefix}neighbors_idx"])


def run(sim_cls, path, exact):
    z, meta, mesh, st, params = load_case(path)
    sim = sim_cls(mesh, meta["real_bytes"], meta["defs"], meta["subs"], capacity=meta["count"], seed=meta["seed"])
    sim.upload(st)
    sim.initialize_mesh()
    sim.step(params, 1)
    check(sim, z, "step1_", meta["real_bytes"], exact)
    sim.step(params, 9)
    check(sim, z, "step10_", meta["real_bytes"], exact)


def test_fixtures_present():
    assert len(GOLDEN) == 4


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_reproduces_golden(path):
    from oracle.pyoracle import OracleSim

    run(OracleSim, path, exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_device_reproduces_golden(path):
    from paraphysicell_amd.mechanics import DeviceSim

    print(f"golden case {os.path.basename(path)}", flush=True)
    run(DeviceSim, path, exact=False)
