"""pcm_neighbors_noalloc, the allocation-free neighbor build: on caller-owned device arrays
its lists equal those of the handle path (pcm_update_mechanics_mesh +
pcm_update_cell_neighbors + pcm_get_neighbors) on the same agents, bit for bit, also on a
second call right after a larger one (the shared scratch area keeps the first call's
contents), in FP32 and FP64, 2D and 3D, and when lists do not fit the row width."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from paraphysicell_amd import abi, synth
from paraphysicell_amd.mechanics import DeviceSim, load_library, neighbors_noalloc

pytestmark = pytest.mark.gpu


def _population(dims, n, seed, real_bytes, immovable_every=0):
    radius = 60.0 * (n / 600) ** (1 / dims) if dims == 3 else 40.0 * (n / 100) ** 0.5
    mesh, st, _ = synth.ball_population(dims, n, radius, 1000, seed, real_bytes, radius_sd=0.5)
    rng = np.random.default_rng(seed)
    st["relative_maximum_adhesion_distance"] = rng.uniform(1.0, 1.6, n)
    if immovable_every:
        mov = np.ones(n, np.uint8)
        mov[::immovable_every] = 0
        st["is_movable"] = mov
    return mesh, st


def _handle_lists(mesh, st, real_bytes):
    sim = DeviceSim(mesh, real_bytes, st.defs, st.subs, capacity=st.count, seed=1)
    sim.upload(st)
    sim.initialize_mesh()
    off, idx = sim.neighbors()
    sim.close()
    return off, idx


class _Hip:
    """Device buffers through the HIP runtime that libpcm.so itself uses (dlopen by soname
    returns the instance already loaded with the library)."""

    def __init__(self):
        load_library()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        self.live.append(p.value)
        if a.nbytes:
            assert self.rt.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        if out.nbytes:
            assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def synchronize(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _noalloc_lists(mesh, st, real_bytes, width):
    hip = _Hip()
    n = st.count
    pos = hip.upload(st["positions"])
    rad = hip.upload(st["radius"])
    rel = hip.upload(st["relative_maximum_adhesion_distance"])
    mov = hip.upload(st["is_movable"].astype(np.uint8))
    counts0 = np.full(max(n, 1), -7, np.int32)
    rows0 = np.full(max(n * width, 1), -7, np.int32)
    needed0 = np.full(1, -7, np.int32)
    counts, rows, needed = hip.upload(counts0), hip.upload(rows0), hip.upload(needed0)
    # the null stream: the kernels follow the uploads above and precede the read-back below
    neighbors_noalloc(mesh, real_bytes, n, pos, rad, rel, mov, width, counts, rows, needed, stream=None)
    hip.synchronize()
    out = (hip.download(counts, counts0)[:n], hip.download(rows, rows0)[: n * width].reshape(n, width),
           int(hip.download(needed, needed0)[0]))
    hip.free()
    return out


def _assert_rows_equal(off, idx, counts, rows, width, what):
    full = np.diff(off)
    assert np.array_equal(counts, np.minimum(full, width)), f"{what}: list lengths"
    for i in np.nonzero(counts)[0]:
        c = counts[i]
        assert np.array_equal(rows[i, :c], idx[off[i]:off[i] + c]), f"{what}: list of agent {i}"


@pytest.mark.parametrize("real_bytes", [4, 8])
def test_noalloc_equals_handle_path_on_two_calls(real_bytes):
    # a larger 3D call, then a smaller 2D one on another grid, then a smaller 3D one
    cases = [(3, 20000, 5, 7), (2, 3000, 6, 0), (3, 800, 7, 5)]
    for dims, n, seed, every in cases:
        what = f"{dims}D, {n} agents, FP{8 * real_bytes}"
        print(f"noalloc vs handle path: {what}", flush=True)
        mesh, st = _population(dims, n, seed, real_bytes, every)
        off, idx = _handle_lists(mesh, st, real_bytes)
        width = int(np.diff(off).max()) + 3
        counts, rows, needed = _noalloc_lists(mesh, st, real_bytes, width)
        assert needed == 0, what
        _assert_rows_equal(off, idx, counts, rows, width, what)
        if every:
            assert not counts[::every].any(), f"{what}: non-movable agents list nobody"


def test_noalloc_rows_too_narrow_report_the_longest_list():
    mesh, st = _population(3, 6000, 9, 4)
    off, idx = _handle_lists(mesh, st, 4)
    longest = int(np.diff(off).max())
    width = max(1, longest // 3)
    print(f"noalloc with width {width} against lists up to {longest}", flush=True)
    counts, rows, needed = _noalloc_lists(mesh, st, 4, width)
    assert needed == longest
    _assert_rows_equal(off, idx, counts, rows, width, "narrow rows")
    # a following call with wide rows is complete again
    counts, rows, needed = _noalloc_lists(mesh, st, 4, longest)
    assert needed == 0
    _assert_rows_equal(off, idx, counts, rows, longest, "wide rows after narrow")


def test_noalloc_empty_population():
    mesh = abi.MeshSpec(3, (-100,) * 3, (100,) * 3)
    from paraphysicell_amd.state import AgentState

    st = AgentState(mesh, 4, 1, 1, 0).defaults()
    counts, rows, needed = _noalloc_lists(mesh, st, 4, 8)
    assert needed == 0 and counts.size == 0
