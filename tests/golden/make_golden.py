"""Writes the golden fixtures of tests/test_golden.py with the oracle (run in the build
container; the reference itself is never run).

Each fixture is a small population (N <= 64) from tests/parity.rich_population with
busy rates, FP32 and FP64, 2D and 3D: the initial state (every field and link list), the
step parameters and seed, and the state after 1 and after 10 fused mechanics steps in
the oracle's keyed-random mode (the mode the device library reproduces).  Files are
.npz without pickles.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle.pyoracle import OracleSim  # noqa: E402
from paraphysicell_amd import abi  # noqa: E402
from parity import rich_population  # noqa: E402

CASES = [(2, 4, 48, 101), (2, 8, 48, 102), (3, 4, 64, 103), (3, 8, 64, 104)]
SEED = 77


def case_name(dims, real_bytes, n, seed):
    return f"rich_{dims}d_{'f32' if real_bytes == 4 else 'f64'}_n{n}_s{seed}"


def snapshot(sim, prefix, out):
    for name, _, _ in abi.FIELDS:
        out[f"{prefix}field_{name}"] = sim.get_field(name)
    for which, label in ((abi.LINKS_SPRINGS, "springs"), (abi.LINKS_ATTACHED, "attached")):
        off, idx = sim.links(which)
        out[f"{prefix}{label}_off"], out[f"{prefix}{label}_idx"] = off, idx
    off, idx = sim.neighbors()
    out[f"{prefix}neighbors_off"], out[f"{prefix}neighbors_idx"] = off, idx
    out[f"{prefix}count"] = np.array([sim.count], np.int64)
    out[f"{prefix}deaths"] = np.array([sim.deaths_count], np.int64)


def main():
    for dims, real_bytes, n, seed in CASES:
        mesh, st, params = rich_population(dims, n, seed, real_bytes)
        out = {}
        for name, _, _ in abi.FIELDS:
            out[f"in_field_{name}"] = st.arrays[name]
        for which, label in ((abi.LINKS_SPRINGS, "springs"), (abi.LINKS_ATTACHED, "attached")):
            out[f"in_{label}_off"], out[f"in_{label}_idx"] = st.links[which]
        if st.gradients is not None:
            out["in_gradient_voxel"] = np.asarray(st.gradients[0], np.int32)
            out["in_gradients"] = st.gradients[1]
        meta = {"dims": dims, "real_bytes": real_bytes, "defs": st.defs, "subs": st.subs, "count": n,
                "seed": SEED, "bbox_min": list(mesh.bbox_min), "bbox_max": list(mesh.bbox_max),
                "voxel": list(mesh.voxel), "dt": params.dt, "virtual_wall": params.virtual_wall,
                "springs": params.springs}
        out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        sim = OracleSim(mesh, real_bytes, st.defs, st.subs, capacity=n, seed=SEED)
        sim.upload(st)
        sim.initialize_mesh()
        sim.step(params, 1)
        snapshot(sim, "step1_", out)
        sim.step(params, 9)
        snapshot(sim, "step10_", out)
        path = os.path.join(HERE, case_name(dims, real_bytes, n, seed) + ".npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
