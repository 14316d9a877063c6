"""The C-ABI library: it loads, exports every entry point include/pcm.h declares, its field
table agrees with the Python mirror, and it refuses to run without a GPU (no fallback)."""
import ctypes
import os
import re
import subprocess

import pytest

from paraphysicell_amd import abi, mechanics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pcm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pcm_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.LIB_PATH):
        from paraphysicell_amd import build

        build.build()
    return mechanics.load_library()


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (pcm_[a-z_0-9]+)", out.stdout))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)


def test_field_table_matches_header(lib):
    t = ctypes.c_int32()
    w = ctypes.c_int32()
    name = ctypes.c_char_p()
    for fid, (pname, ptype, pwidth) in enumerate(abi.FIELDS):
        assert lib.pcm_field_info(fid, ctypes.byref(t), ctypes.byref(w), ctypes.byref(name)) == abi.PCM_OK
        assert (t.value, w.value) == (ptype, pwidth), pname
        assert name.value.decode().lower() == pname.lower(), (name.value, pname)
    assert lib.pcm_field_info(len(abi.FIELDS), None, None, None) == abi.PCM_ERR_ARGUMENT


def test_version(lib):
    assert lib.pcm_version() >= 1


def test_no_silent_fallback(lib):
    """Without a usable GPU the library reports an error instead of computing on the CPU."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(mechanics.PcmError) as e:
        mechanics.DeviceSim(abi.MeshSpec(3, (0, 0, 0), (90, 90, 90)), 4, 1, 1, capacity=8)
    assert e.value.code == abi.PCM_ERR_DEVICE


def test_argument_errors(lib):
    h = ctypes.c_void_p()
    assert lib.pcm_create(None, ctypes.byref(h)) == abi.PCM_ERR_ARGUMENT
    cfg = abi.make_config(abi.MeshSpec(4, (0, 0, 0), (9, 9, 9)), 4, 1, 1, 8, 0)
    assert lib.pcm_create(ctypes.byref(cfg), ctypes.byref(h)) == abi.PCM_ERR_ARGUMENT
    cfg = abi.make_config(abi.MeshSpec(3, (0, 0, 0), (9, 9, 9)), 2, 1, 1, 8, 0)
    assert lib.pcm_create(ctypes.byref(cfg), ctypes.byref(h)) in (abi.PCM_ERR_ARGUMENT, abi.PCM_ERR_DEVICE)
    assert lib.pcm_step(None, None, 1) == abi.PCM_ERR_ARGUMENT


def test_batch_argument_errors(lib):
    b = ctypes.c_void_p()
    assert lib.pcm_batch_create(3, ctypes.byref(b)) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_create(4, ctypes.byref(b)) == abi.PCM_OK
    n = ctypes.c_int32(-1)
    assert lib.pcm_batch_size(b, ctypes.byref(n)) == abi.PCM_OK and n.value == 0
    assert lib.pcm_batch_admit(b, None) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_retire(b, None) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_step(b, None, -1) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_step(b, None, 1) == abi.PCM_OK  # no members: nothing to do
    assert lib.pcm_batch_step(None, None, 1) == abi.PCM_ERR_ARGUMENT
    out = (ctypes.c_int32 * 1)()
    assert lib.pcm_batch_counts(None, 0, None, None) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_counts(b, 1, None, out) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_counts(b, 0, None, None) == abi.PCM_OK  # nothing asked
    assert lib.pcm_batch_get_tuning(b, 3, None) == abi.PCM_ERR_ARGUMENT
    t = abi.Tuning(8, 8, 2)
    assert lib.pcm_batch_set_tuning(None, 3, ctypes.byref(t)) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_get_fields(None, 0, None, None) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_destroy(b) == abi.PCM_OK


def test_tuning_argument_errors(lib):
    t = abi.Tuning(8, 8, 2)
    assert lib.pcm_set_tuning(None, ctypes.byref(t)) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_autotune(None, None, 1, None) == abi.PCM_ERR_ARGUMENT
    assert lib.pcm_batch_autotune(None, None, 1) == abi.PCM_ERR_ARGUMENT


def test_noalloc_refuses_bad_sizes_before_touching_the_device(lib):
    """The allocation-free neighbor build checks its arguments and its scratch need on the
    host, before any device call, so these errors come back on a machine without a GPU too."""
    fake = 4096  # never dereferenced: every call below fails its host-side checks first
    small = abi.MeshSpec(3, (0, 0, 0), (90, 90, 90))
    with pytest.raises(mechanics.PcmError) as e:
        mechanics.neighbors_noalloc(small, 2, 10, fake, fake, fake, None, 8, fake, fake, fake, lib=lib)
    assert e.value.code == abi.PCM_ERR_ARGUMENT
    with pytest.raises(mechanics.PcmError) as e:
        mechanics.neighbors_noalloc(small, 4, 10, None, fake, fake, None, 8, fake, fake, fake, lib=lib)
    assert e.value.code == abi.PCM_ERR_ARGUMENT
    with pytest.raises(mechanics.PcmError) as e:
        mechanics.neighbors_noalloc(small, 4, 10, fake, fake, fake, None, 0, fake, fake, fake, lib=lib)
    assert e.value.code == abi.PCM_ERR_ARGUMENT
    # 334^3 = 37M voxels need 298 MB of voxel counts and starts: beyond the fixed area
    huge = abi.MeshSpec(3, (-5000,) * 3, (5000,) * 3)
    with pytest.raises(mechanics.PcmError) as e:
        mechanics.neighbors_noalloc(huge, 4, 10, fake, fake, fake, None, 8, fake, fake, fake, lib=lib)
    assert e.value.code == abi.PCM_ERR_CAPACITY
    assert "PCM_NOALLOC_SCRATCH_BYTES" in str(e.value)
    # so do 5M agents on a small grid
    with pytest.raises(mechanics.PcmError) as e:
        mechanics.neighbors_noalloc(small, 8, 5_000_000, fake, fake, fake, None, 1, fake, fake, fake, lib=lib)
    assert e.value.code == abi.PCM_ERR_CAPACITY


def test_mesh_lookups_follow_the_device_binning(lib):
    """pcm_get_mesh_position / pcm_get_mesh_index (common_solver::get_mesh_position /
    get_mesh_index) need no GPU: voxel coordinates by (int)((p - min) / voxel) in real_t,
    clamped to the grid (ceil of extent / voxel), linearised x fastest."""
    import numpy as np

    lib.pcm_get_mesh_position.argtypes = [ctypes.POINTER(abi.MeshDesc), ctypes.c_int32, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_int32 * 3)]
    lib.pcm_get_mesh_index.argtypes = [ctypes.POINTER(abi.MeshDesc), ctypes.POINTER(ctypes.c_int32 * 3),
                                       ctypes.POINTER(ctypes.c_int32)]
    rng = np.random.default_rng(3)
    for dims in (1, 2, 3):
        mesh = abi.MeshSpec(dims, (-100, -70, -40), (101, 80, 45), (30, 20, 15))
        desc = mesh.to_c()
        g = mesh.grid_shape()
        for real in (np.float32, np.float64):
            pts = rng.uniform(-130, 130, size=(200, dims)).astype(real)
            pts[:3] = np.array([-100, -70, -40][:dims], real)  # the box minimum
            pts[3:6] = np.array([101, 80, 45][:dims], real)  # the box maximum: clamped into the last voxel
            for p in pts:
                out = (ctypes.c_int32 * 3)()
                a = np.ascontiguousarray(p)
                assert lib.pcm_get_mesh_position(ctypes.byref(desc), a.itemsize, a.ctypes.data, ctypes.byref(out)) == 0
                want = [0, 0, 0]
                for k in range(dims):
                    q = (p[k] - real(mesh.bbox_min[k])) / real(mesh.voxel[k])
                    want[k] = int(min(max(int(q), 0), g[k] - 1))
                assert list(out) == want, (dims, real, p)
                idx = ctypes.c_int32()
                assert lib.pcm_get_mesh_index(ctypes.byref(desc), ctypes.byref(out), ctypes.byref(idx)) == 0
                assert idx.value == want[0] + g[0] * (want[1] + g[1] * want[2])
        bad = (ctypes.c_int32 * 3)(g[0], 0, 0)
        assert lib.pcm_get_mesh_index(ctypes.byref(desc), ctypes.byref(bad), ctypes.byref(ctypes.c_int32())) \
            == abi.PCM_ERR_ARGUMENT


@pytest.mark.gpu
def test_cell_type_outside_the_definitions_is_refused():
    """A cell_definition_index outside [0, cell_defs) would index the per-type rows out of
    bounds on the device: pcm_set_field refuses it."""
    import numpy as np

    sim = mechanics.DeviceSim(abi.MeshSpec(3, (0, 0, 0), (90, 90, 90)), 4, 2, 1, capacity=8)
    sim._call("set_count", sim._h, 4)
    for bad in ([0, 1, 2, 0], [0, -1, 0, 0]):
        a = np.array(bad, np.int32)
        with pytest.raises(mechanics.PcmError) as e:
            sim._call("set_field", sim._h, abi.FIELD_ID["cell_definition_index"], a.ctypes.data)
        assert e.value.code == abi.PCM_ERR_ARGUMENT
    ok = np.array([0, 1, 1, 0], np.int32)
    sim._call("set_field", sim._h, abi.FIELD_ID["cell_definition_index"], ok.ctypes.data)
    sim.close()
