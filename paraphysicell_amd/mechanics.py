"""Python host side of the MI355X mechanics library (libpcm.so, include/pcm.h).

``DeviceSim`` owns one simulation's device-resident state.  ``Solver`` mirrors the
reference's solver aggregate (src/solver/solver.h:9-17): ``solver.position``,
``solver.containers`` and ``solver.interactions`` expose the reference's entry points by
name, each one forwarding to the matching C-ABI call.

There is no CPU fallback: constructing a ``DeviceSim`` without the built library, or on
a machine without a GPU, raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import abi
from .state import AgentState


class PcmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"pcm error {code}: {message}")
        self.code = code


_PTR = ctypes.c_void_p


def _declare(lib, prefix: str):
    """Argument types of the C entry points shared by libpcm.so and the oracle."""
    P = _PTR
    sig = {
        "create": [ctypes.POINTER(abi.Config), ctypes.POINTER(P)],
        "destroy": [P],
        "set_count": [P, ctypes.c_int32],
        "get_count": [P, ctypes.POINTER(ctypes.c_int32)],
        "set_field": [P, ctypes.c_int32, P],
        "get_field": [P, ctypes.c_int32, P],
        "set_links": [P, ctypes.c_int32, P, P],
        "get_links_size": [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)],
        "get_links": [P, ctypes.c_int32, P, P],
        "set_neighbors": [P, P, P],
        "get_neighbors_size": [P, ctypes.POINTER(ctypes.c_int64)],
        "get_neighbors": [P, P, P],
        "get_voxel_count": [P, ctypes.POINTER(ctypes.c_int64)],
        "get_voxel_buckets": [P, P, P],
        "get_deaths_count": [P, ctypes.POINTER(ctypes.c_int64)],
        "set_gradients": [P, P, P],
        "set_step_index": [P, ctypes.c_uint64],
        "get_step_index": [P, ctypes.POINTER(ctypes.c_uint64)],
        "update_mechanics_mesh": None,
        "update_cell_neighbors": None,
        "evaluate_interactions": [P, ctypes.POINTER(abi.StepParams)],
        "update_cell_forces": [P, ctypes.POINTER(abi.StepParams)],
        "update_motility": [P, ctypes.POINTER(abi.StepParams)],
        "update_basement_membrane_interactions": [P, ctypes.POINTER(abi.StepParams)],
        "update_spring_attachments": [P, ctypes.POINTER(abi.StepParams)],
        "update_positions": [P, ctypes.POINTER(abi.StepParams)],
        "update_cell_cell_interactions": [P, ctypes.POINTER(abi.StepParams)],
        "update_cell_container_for_mechanics": [P, ctypes.POINTER(abi.StepParams)],
        "update_cell_container_for_phenotype": [P, ctypes.POINTER(abi.StepParams), ctypes.c_int32],
        "get_divisions_count": [P, ctypes.POINTER(ctypes.c_int64)],
        "division_log": [P, P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)],
        "step": [P, ctypes.POINTER(abi.StepParams), ctypes.c_int32],
        "set_step_params": [P, ctypes.POINTER(abi.StepParams)],
    }
    for name, args in sig.items():
        fn = getattr(lib, prefix + name)
        fn.restype = ctypes.c_int
        if args is not None:
            fn.argtypes = args
    getattr(lib, prefix + "last_error").restype = ctypes.c_char_p


class _LibSim:
    """Common wrapper over a C library exposing the pcm.h entry points under ``prefix``."""

    prefix = "pcm_"
    _mesh_only_phases = ("update_mechanics_mesh", "update_cell_neighbors")

    def __init__(self, lib, mesh: abi.MeshSpec, real_bytes: int = 4, defs: int = 1, subs: int = 1,
                 capacity: int = 1024, seed: int = 0):
        self._lib = lib
        self.mesh = mesh
        self.real_bytes = real_bytes
        self.defs = defs
        self.subs = subs
        self.capacity = capacity
        self.seed = seed
        cfg = abi.make_config(mesh, real_bytes, defs, subs, capacity, seed)
        h = _PTR()
        self._h = None
        self._call("create", ctypes.byref(cfg), ctypes.byref(h))
        self._h = h

    # ------------------------------------------------------------------ plumbing
    def _fn(self, name):
        return getattr(self._lib, self.prefix + name)

    def _call(self, name, *args):
        rc = self._fn(name)(*args)
        if rc != abi.PCM_OK:
            msg = self._fn("last_error")()
            raise PcmError(rc, msg.decode() if msg else "")

    def close(self):
        if self._h is not None:
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ data movement
    @property
    def count(self) -> int:
        n = ctypes.c_int32()
        self._call("get_count", self._h, ctypes.byref(n))
        return n.value

    def upload(self, state: AgentState, fields=None) -> None:
        if state.count > self.capacity:
            raise PcmError(abi.PCM_ERR_CAPACITY, "state larger than capacity")
        self._call("set_count", self._h, state.count)
        arrays = [(abi.FIELD_ID[name], np.ascontiguousarray(state.arrays[name])) for name, _, _ in abi.FIELDS
                  if fields is None or name in fields]
        bulk = getattr(self._lib, self.prefix + "set_fields", None)
        if bulk is not None and arrays:
            # every field in one call: one count read, one wait (pcm_set_fields)
            ids = (ctypes.c_int32 * len(arrays))(*[fid for fid, _ in arrays])
            ptrs = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for _, a in arrays])
            bulk.argtypes = [_PTR, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
            bulk.restype = ctypes.c_int
            rc = bulk(self._h, len(arrays), ids, ptrs)
            if rc != abi.PCM_OK:
                msg = self._fn("last_error")()
                raise PcmError(rc, msg.decode() if msg else "")
        else:
            for fid, a in arrays:
                self._call("set_field", self._h, fid, a.ctypes.data)
        for which in (abi.LINKS_SPRINGS, abi.LINKS_ATTACHED):
            off, idx = state.links[which]
            off = np.ascontiguousarray(off, dtype=np.int32)
            idx = np.ascontiguousarray(idx, dtype=np.int32)
            if idx.size == 0:
                idx = np.zeros(1, np.int32)
            self._call("set_links", self._h, which, off.ctypes.data, idx.ctypes.data)
        if state.neighbors is not None:
            off = np.ascontiguousarray(state.neighbors[0], dtype=np.int32)
            idx = np.ascontiguousarray(state.neighbors[1], dtype=np.int32)
            if idx.size == 0:
                idx = np.zeros(1, np.int32)
            self._call("set_neighbors", self._h, off.ctypes.data, idx.ctypes.data)
        if state.gradients is not None:
            self.set_gradients(*state.gradients)

    def set_gradients(self, diffusion_voxel, gradients) -> None:
        """Substrate gradients per diffusion voxel (voxels x substrates x dims, x fastest
        over voxels), the input of the chemotaxis direction updates."""
        vox = np.ascontiguousarray(diffusion_voxel, dtype=np.int32).reshape(3)
        g = np.ascontiguousarray(gradients, dtype=np.float32 if self.real_bytes == 4 else np.float64)
        voxels = int(np.prod(self.mesh.with_voxel(tuple(int(v) for v in vox)).grid_shape()))
        need = voxels * self.subs * self.mesh.dims
        if g.size != need:
            raise PcmError(abi.PCM_ERR_ARGUMENT, f"gradients: {g.size} values, the diffusion grid needs {need}")
        self._call("set_gradients", self._h, vox.ctypes.data, g.ctypes.data)

    def get_field(self, name: str) -> np.ndarray:
        ftype, wkind = abi.FIELDS[abi.FIELD_ID[name]][1:]
        n = self.count
        w = abi.field_width(wkind, self.mesh.dims, self.defs, self.subs)
        out = np.zeros(max(n * w, 1), dtype=abi.field_dtype(ftype, self.real_bytes))
        self._call("get_field", self._h, abi.FIELD_ID[name], out.ctypes.data)
        return out[: n * w]

    def get_fields(self, names) -> dict:
        """Several fields in one call (pcm_get_fields: one count read, one wait); field by
        field where the library has no bulk read."""
        bulk = getattr(self._lib, self.prefix + "get_fields", None)
        if bulk is None:
            return {name: self.get_field(name) for name in names}
        n = self.count
        out, widths = {}, {}
        for name in names:
            ftype, wkind = abi.FIELDS[abi.FIELD_ID[name]][1:]
            widths[name] = abi.field_width(wkind, self.mesh.dims, self.defs, self.subs)
            out[name] = np.zeros(max(n * widths[name], 1), dtype=abi.field_dtype(ftype, self.real_bytes))
        if out:
            ids = (ctypes.c_int32 * len(out))(*[abi.FIELD_ID[name] for name in out])
            ptrs = (ctypes.c_void_p * len(out))(*[a.ctypes.data for a in out.values()])
            bulk.argtypes = [_PTR, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
            bulk.restype = ctypes.c_int
            rc = bulk(self._h, len(out), ids, ptrs)
            if rc != abi.PCM_OK:
                msg = self._fn("last_error")()
                raise PcmError(rc, msg.decode() if msg else "")
        return {name: a[: n * widths[name]] for name, a in out.items()}

    def download(self) -> AgentState:
        n = self.count
        st = AgentState(self.mesh, self.real_bytes, self.defs, self.subs, n)
        for name, _, _ in abi.FIELDS:
            st.arrays[name] = self.get_field(name)
        for which in (abi.LINKS_SPRINGS, abi.LINKS_ATTACHED):
            st.links[which] = self.links(which)
        st.neighbors = self.neighbors()
        return st

    def links(self, which: int):
        total = ctypes.c_int64()
        self._call("get_links_size", self._h, which, ctypes.byref(total))
        off = np.zeros(self.count + 1, np.int32)
        idx = np.zeros(max(total.value, 1), np.int32)
        self._call("get_links", self._h, which, off.ctypes.data, idx.ctypes.data)
        return off, idx[: total.value]

    def neighbors(self):
        total = ctypes.c_int64()
        self._call("get_neighbors_size", self._h, ctypes.byref(total))
        off = np.zeros(self.count + 1, np.int32)
        idx = np.zeros(max(total.value, 1), np.int32)
        self._call("get_neighbors", self._h, off.ctypes.data, idx.ctypes.data)
        return off, idx[: total.value]

    def buckets(self):
        v = ctypes.c_int64()
        self._call("get_voxel_count", self._h, ctypes.byref(v))
        counts = np.zeros(v.value, np.int32)
        agents = np.zeros(max(self.capacity, self.count, 1), np.int32)  # last binning may predate removals
        self._call("get_voxel_buckets", self._h, counts.ctypes.data, agents.ctypes.data)
        return counts, agents[: int(counts.sum())]

    @property
    def deaths_count(self) -> int:
        d = ctypes.c_int64()
        self._call("get_deaths_count", self._h, ctypes.byref(d))
        return d.value

    @property
    def divisions_count(self) -> int:
        d = ctypes.c_int64()
        self._call("get_divisions_count", self._h, ctypes.byref(d))
        return d.value

    def division_log(self) -> np.ndarray:
        """Parents of the last phenotype update's divisions, in division order."""
        length = ctypes.c_int64()
        self._call("division_log", self._h, None, 0, ctypes.byref(length))
        out = np.zeros(max(length.value, 1), np.int32)
        self._call("division_log", self._h, out.ctypes.data, length.value, ctypes.byref(length))
        return out[: length.value]

    @property
    def step_index(self) -> int:
        s = ctypes.c_uint64()
        self._call("get_step_index", self._h, ctypes.byref(s))
        return s.value

    @step_index.setter
    def step_index(self, value: int) -> None:
        self._call("set_step_index", self._h, int(value))

    # ------------------------------------------------------------------ phases
    def _phase(self, name, params):
        if name in self._mesh_only_phases:
            self._call(name, self._h)
        else:
            p = (params or abi.Params()).to_c()
            self._call(name, self._h, ctypes.byref(p))

    def update_mechanics_mesh(self, params=None):
        self._phase("update_mechanics_mesh", params)

    def update_cell_neighbors(self, params=None):
        self._phase("update_cell_neighbors", params)

    def evaluate_interactions(self, params=None):
        """Contact functions over the attached cells (src/simulator.cpp:65-82)."""
        self._phase("evaluate_interactions", params)

    def update_cell_forces(self, params=None):
        self._phase("update_cell_forces", params)

    def update_motility(self, params=None):
        self._phase("update_motility", params)

    def update_basement_membrane_interactions(self, params=None):
        self._phase("update_basement_membrane_interactions", params)

    def update_spring_attachments(self, params=None):
        self._phase("update_spring_attachments", params)

    def update_positions(self, params=None):
        self._phase("update_positions", params)

    def update_cell_cell_interactions(self, params=None):
        self._phase("update_cell_cell_interactions", params)

    def update_cell_container_for_mechanics(self, params=None):
        self._phase("update_cell_container_for_mechanics", params)

    def update_cell_container_for_phenotype(self, params=None, compute_internalized_substrates=False):
        """Divisions of agents flagged to divide, then the removal loop
        (src/solver/host/containers_solver.cpp:149-187)."""
        p = (params or abi.Params()).to_c()
        self._call("update_cell_container_for_phenotype", self._h, ctypes.byref(p),
                   1 if compute_internalized_substrates else 0)

    def initialize_mesh(self):
        """simulator::initialize tail (src/simulator.cpp:47-48): mesh, then neighbors."""
        self.update_mechanics_mesh()
        self.update_cell_neighbors()

    def step(self, params=None, steps: int = 1):
        p = (params or abi.Params()).to_c()
        self._call("step", self._h, ctypes.byref(p), int(steps))

    def set_step_params(self, params) -> None:
        """The parameters this simulation steps under in a batch stepped with params None."""
        p = params.to_c()
        self._call("set_step_params", self._h, ctypes.byref(p))

    def run_phases(self, params=None):
        """The mechanics step as the sequence of per-phase calls (src/simulator.cpp:171-192)."""
        for name in ("evaluate_interactions", "update_cell_forces", "update_motility", "update_basement_membrane_interactions",
                     "update_spring_attachments", "update_positions", "update_cell_cell_interactions",
                     "update_cell_container_for_mechanics", "update_mechanics_mesh", "update_cell_neighbors"):
            getattr(self, name)(params)


_LIB = None


def load_library(path: str | None = None):
    """Load libpcm.so; raises if it has not been built (no fallback exists)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or abi.LIB_PATH
    if not os.path.exists(p):
        raise PcmError(abi.PCM_ERR_UNSUPPORTED,
                       f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(p)
    _declare(lib, "pcm_")
    lib.pcm_synchronize.argtypes = [_PTR]
    lib.pcm_synchronize.restype = ctypes.c_int
    lib.pcm_field_info.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_char_p)]
    lib.pcm_field_info.restype = ctypes.c_int
    lib.pcm_version.restype = ctypes.c_int
    lib.pcm_kernel_name.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p)]
    lib.pcm_kernel_name.restype = ctypes.c_int
    lib.pcm_profile_begin.argtypes = [_PTR, ctypes.c_uint64]
    lib.pcm_profile_begin.restype = ctypes.c_int
    lib.pcm_profile_end.argtypes = [_PTR, ctypes.c_void_p, ctypes.c_void_p]
    lib.pcm_profile_end.restype = ctypes.c_int
    lib.pcm_get_stats.argtypes = [_PTR, ctypes.c_void_p]
    lib.pcm_get_stats.restype = ctypes.c_int
    for name, args in (("batch_create", [ctypes.c_int32, ctypes.POINTER(_PTR)]), ("batch_destroy", [_PTR]),
                       ("batch_admit", [_PTR, _PTR]), ("batch_retire", [_PTR, _PTR]),
                       ("batch_size", [_PTR, ctypes.POINTER(ctypes.c_int32)]),
                       ("batch_step", [_PTR, ctypes.POINTER(abi.StepParams), ctypes.c_int32]),
                       ("batch_synchronize", [_PTR]), ("batch_use_graphs", [_PTR, ctypes.c_int32]),
                       ("batch_flatten", [_PTR, ctypes.c_int32]),
                       ("batch_autotune", [_PTR, ctypes.POINTER(abi.StepParams), ctypes.c_int32]),
                       ("set_tuning", [_PTR, ctypes.POINTER(abi.Tuning)]),
                       ("get_tuning", [_PTR, ctypes.POINTER(abi.Tuning)]),
                       ("autotune", [_PTR, ctypes.POINTER(abi.StepParams), ctypes.c_int32,
                                     ctypes.POINTER(abi.Tuning)])):
        fn = getattr(lib, "pcm_" + name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    if path is None:
        _LIB = lib
    return lib


class DeviceSim(_LibSim):
    """One simulation's mechanics state resident on the GPU (libpcm.so)."""

    prefix = "pcm_"

    def __init__(self, mesh: abi.MeshSpec, real_bytes: int = 4, defs: int = 1, subs: int = 1,
                 capacity: int = 1024, seed: int = 0, lib=None):
        super().__init__(lib or load_library(), mesh, real_bytes, defs, subs, capacity, seed)

    def synchronize(self):
        self._call("synchronize", self._h)

    def removal_log(self) -> np.ndarray:
        """Positions passed to the reference's remove_at since the last read, in order
        (pcm_removal_log); reading restarts the log."""
        fn = self._fn("removal_log")
        fn.argtypes = [_PTR, _PTR, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        fn.restype = ctypes.c_int
        length = ctypes.c_int64()
        self._call("removal_log", self._h, None, 0, ctypes.byref(length))
        out = np.zeros(max(length.value, 1), np.int32)
        self._call("removal_log", self._h, out.ctypes.data, length.value, ctypes.byref(length))
        return out[: length.value]

    def pack_fields(self, names, widths=None) -> np.ndarray:
        """The named fields of every agent as one (count, columns) float64 matrix packed on
        the device (pcm_pack_fields); widths pad vectors (0: the field's own width)."""
        widths = [0] * len(names) if widths is None else list(widths)
        fields = np.array([abi.FIELD_ID[n] for n in names], np.int32)
        w = np.array(widths, np.int32)
        row = 0
        for name, wk in zip(names, widths):
            own = abi.field_width(abi.FIELDS[abi.FIELD_ID[name]][2], self.mesh.dims, self.defs, self.subs)
            row += wk or own
        n = self.count
        out = np.zeros(max(n * row, 1), np.float64)
        fn = self._fn("pack_fields")
        fn.argtypes = [_PTR, ctypes.c_int32, _PTR, _PTR, _PTR]
        fn.restype = ctypes.c_int
        if n > 0:
            self._call("pack_fields", self._h, len(fields), fields.ctypes.data, w.ctypes.data, out.ctypes.data)
        return out[: n * row].reshape(n, row)

    # ------------------------------------------------------------------ launch shape
    @property
    def tuning(self) -> tuple:
        t = abi.Tuning()
        self._call("get_tuning", self._h, ctypes.byref(t))
        return t.as_tuple()

    @tuning.setter
    def tuning(self, value) -> None:
        t = abi.Tuning(*value)
        self._call("set_tuning", self._h, ctypes.byref(t))

    def autotune(self, params=None, steps_per_candidate: int = 5) -> tuple:
        """Pick the launch shape on this simulation's own steps (advances 1 + 6k steps)."""
        p = (params or abi.Params()).to_c()
        t = abi.Tuning()
        self._call("autotune", self._h, ctypes.byref(p), int(steps_per_candidate), ctypes.byref(t))
        return t.as_tuple()

    STAT_NAMES = ("attach_events", "resolve_rounds", "attaching_agents", "interaction_candidates",
                  "interaction_replay", "removed", "occupied_voxels", "reserved")

    def stats(self) -> dict:
        """Work counters of the most recent step."""
        out = (ctypes.c_int32 * 8)()
        self._call("get_stats", self._h, out)
        return dict(zip(self.STAT_NAMES, list(out)))

    # ------------------------------------------------------------------ kernel timing
    def profile_begin(self, kernels) -> None:
        """Bracket every launch of the named kernels (see kernel_names()) with HIP events."""
        names = kernel_names(self._lib)
        mask = 0
        for k in kernels:
            mask |= 1 << names.index(k)
        self._call("profile_begin", self._h, ctypes.c_uint64(mask))

    def profile_every(self, every: int) -> None:
        """Time one launch in `every` of each profiled kernel."""
        fn = self._fn("profile_every")
        fn.argtypes = [_PTR, ctypes.c_int32]
        fn.restype = ctypes.c_int
        self._call("profile_every", self._h, int(every))

    def profile_end(self) -> dict:
        """{kernel name: (total ms, launches)} for the kernels timed since profile_begin."""
        names = kernel_names(self._lib)
        ms = (ctypes.c_double * len(names))()
        cnt = (ctypes.c_int64 * len(names))()
        self._call("profile_end", self._h, ms, cnt)
        return {names[k]: (ms[k], cnt[k]) for k in range(len(names)) if cnt[k] > 0}


class Batch:
    """Simulations advanced together (pcm_batch_*): one launch per kernel covers every
    member of a dimensionality, whatever its size; results equal stepping each alone."""

    def __init__(self, real_bytes: int = 4, graphs: bool = False, lib=None):
        self._lib = lib or load_library()
        self._h = _PTR()
        self._check(self._lib.pcm_batch_create(int(real_bytes), ctypes.byref(self._h)))
        self._check(self._lib.pcm_batch_use_graphs(self._h, 1 if graphs else 0))
        self.members: list = []

    def _check(self, rc):
        if rc != abi.PCM_OK:
            raise PcmError(rc, self._lib.pcm_last_error().decode())

    def admit(self, sim: "DeviceSim") -> None:
        self._check(self._lib.pcm_batch_admit(self._h, sim._h))
        self.members.append(sim)

    def retire(self, sim: "DeviceSim") -> None:
        self._check(self._lib.pcm_batch_retire(self._h, sim._h))
        self.members.remove(sim)

    @property
    def size(self) -> int:
        n = ctypes.c_int32()
        self._check(self._lib.pcm_batch_size(self._h, ctypes.byref(n)))
        return n.value

    def step(self, params=None, steps: int = 1) -> None:
        """Advance every member `steps` steps: under `params`, or with params None each
        member under its own (DeviceSim.set_step_params or its last step / phase call)."""
        if params is None:
            self._check(self._lib.pcm_batch_step(self._h, None, int(steps)))
            return
        p = params.to_c()
        self._check(self._lib.pcm_batch_step(self._h, ctypes.byref(p), int(steps)))

    def synchronize(self) -> None:
        self._check(self._lib.pcm_batch_synchronize(self._h))

    def use_graphs(self, on: bool) -> None:
        self._check(self._lib.pcm_batch_use_graphs(self._h, 1 if on else 0))

    def use_flattened_grids(self, on: bool) -> None:
        """One launch over exactly the members' blocks (default) or one block row per member."""
        self._check(self._lib.pcm_batch_flatten(self._h, 1 if on else 0))

    def autotune(self, params=None, steps_per_candidate: int = 5) -> None:
        """Pick each dimensionality group's launch shape on the batch's own steps."""
        p = None if params is None else ctypes.byref(params.to_c())
        self._check(self._lib.pcm_batch_autotune(self._h, p, int(steps_per_candidate)))

    def tuning(self, dims: int) -> tuple:
        """The launch shape of the `dims` group (velocity lanes, event lanes, neighbor waves)."""
        t = abi.Tuning()
        self._check(self._lib.pcm_batch_get_tuning(self._h, int(dims), ctypes.byref(t)))
        return t.as_tuple()

    def set_tuning(self, dims: int, value) -> None:
        """Launch shape of the `dims` group, kept for members that join later."""
        t = abi.Tuning(*value)
        self._check(self._lib.pcm_batch_set_tuning(self._h, int(dims), ctypes.byref(t)))

    def counts(self, sims) -> list[int]:
        """Agent counts of member simulations as the last step call left them (the scalars
        the batch gathered then; no device read, unlike DeviceSim.count)."""
        sims = list(sims)
        if not sims:
            return []
        handles = (_PTR * len(sims))(*[s._h for s in sims])
        out = (ctypes.c_int32 * len(sims))()
        self._check(self._lib.pcm_batch_counts(self._h, len(sims), handles, out))
        return list(out)

    def close(self) -> None:
        if self._h:
            self._lib.pcm_batch_destroy(self._h)
            self._h = _PTR()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kernel_names(lib=None):
    lib = lib or load_library()
    out = []
    name = ctypes.c_char_p()
    k = 0
    while lib.pcm_kernel_name(k, ctypes.byref(name)) == abi.PCM_OK:
        out.append(name.value.decode())
        k += 1
    return out


class _Facade:
    def __init__(self, sim, names):
        self._sim = sim
        for n in names:
            setattr(self, n, getattr(sim, n))


class Solver:
    """Reference-shaped aggregate: solver.position / .containers / .interactions."""

    def __init__(self, sim: _LibSim):
        self.sim = sim
        self.position = _Facade(sim, ["update_cell_forces", "update_cell_neighbors", "update_motility",
                                      "update_basement_membrane_interactions", "update_spring_attachments",
                                      "update_positions"])
        self.containers = _Facade(sim, ["update_mechanics_mesh", "update_cell_container_for_mechanics",
                                        "update_cell_container_for_phenotype"])
        self.interactions = _Facade(sim, ["update_cell_cell_interactions"])
        # the free function evaluate_interactions(environment&) of src/simulator.cpp:65-82
        self.evaluate_interactions = sim.evaluate_interactions

    def initialize(self):
        self.sim.initialize_mesh()


def neighbors_noalloc(mesh: abi.MeshSpec, real_bytes: int, count: int, positions, radius, rel_adh, movable,
                      max_neighbors: int, counts, indices, needed, stream=None, lib=None) -> None:
    """pcm_neighbors_noalloc: the mechanics mesh and neighbor lists of `count` agents whose
    fields already sit in device memory, with no device allocation and no host round trip.

    Buffers are device addresses (ints) or objects with ``data_ptr()`` (torch tensors):
    positions (count * dims reals), radius and rel_adh (count reals), movable (count bytes
    or None); outputs counts (count int32), indices (count * max_neighbors int32) and
    needed (one int32: 0, or the longest list when some list did not fit).  The kernels run
    on `stream` (a raw hipStream_t address, or an object with ``cuda_stream``); the call
    returns without waiting for them."""
    lib = lib or load_library()
    fn = lib.pcm_neighbors_noalloc
    fn.argtypes = [ctypes.POINTER(abi.MeshDesc), ctypes.c_int32, ctypes.c_int32, _PTR, _PTR, _PTR, _PTR,
                   ctypes.c_int32, _PTR, _PTR, _PTR, _PTR]
    fn.restype = ctypes.c_int

    def addr(b):
        if b is None:
            return None
        return b.data_ptr() if hasattr(b, "data_ptr") else int(b)

    s = getattr(stream, "cuda_stream", stream)
    desc = mesh.to_c()
    rc = fn(ctypes.byref(desc), real_bytes, count, addr(positions), addr(radius), addr(rel_adh), addr(movable),
            max_neighbors, addr(counts), addr(indices), addr(needed), s)
    if rc != abi.PCM_OK:
        msg = lib.pcm_last_error()
        raise PcmError(rc, msg.decode() if msg else "")
