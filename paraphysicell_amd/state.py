"""Host-side agent state in the reference's SoA layout (src/cell_data.h:14-196).

One numpy array per field (``count * width`` elements, agent-major) plus CSR link lists
for springs and attached cells.  ``AgentState.defaults`` fills the phenotype defaults of
src/data_storage_phenotype.cpp:62-328 and src/cell.cpp:47-60.
"""
from __future__ import annotations

import copy

import numpy as np

from . import abi


class AgentState:
    def __init__(self, mesh: abi.MeshSpec, real_bytes: int, defs: int, subs: int, count: int):
        self.mesh = mesh
        self.real_bytes = real_bytes
        self.defs = defs
        self.subs = subs
        self.count = int(count)
        self.arrays: dict[str, np.ndarray] = {}
        for name, ftype, wkind in abi.FIELDS:
            w = abi.field_width(wkind, mesh.dims, defs, subs)
            self.arrays[name] = np.zeros(self.count * w, dtype=abi.field_dtype(ftype, real_bytes))
        self.arrays["origin"][:] = np.arange(self.count, dtype=np.int32)
        empty = (np.zeros(self.count + 1, np.int32), np.zeros(0, np.int32))
        self.links = {abi.LINKS_SPRINGS: empty, abi.LINKS_ATTACHED: (empty[0].copy(), empty[1].copy())}
        # states.neighbors as CSR, or None when the state carries no lists
        self.neighbors = None
        # (diffusion voxel shape, substrate gradients) for the chemotaxis models, or None
        self.gradients = None

    # ------------------------------------------------------------------ access
    @property
    def dims(self) -> int:
        return self.mesh.dims

    @property
    def real(self):
        return np.float32 if self.real_bytes == 4 else np.float64

    def width(self, name: str) -> int:
        _, _, wkind = abi.FIELDS[abi.FIELD_ID[name]]
        return abi.field_width(wkind, self.dims, self.defs, self.subs)

    def __getitem__(self, name: str) -> np.ndarray:
        return self.arrays[name]

    def __setitem__(self, name: str, value) -> None:
        a = self.arrays[name]
        v = np.asarray(value, dtype=a.dtype)
        a[...] = v if v.size == 1 else v.reshape(a.shape)

    def view(self, name: str) -> np.ndarray:
        """Field as a (count, width) array sharing memory."""
        return self.arrays[name].reshape(self.count, self.width(name))

    def copy(self) -> "AgentState":
        return copy.deepcopy(self)

    def set_links(self, which: int, lists) -> None:
        lens = np.array([len(l) for l in lists], dtype=np.int64)
        off = np.zeros(self.count + 1, np.int32)
        off[1:] = np.cumsum(lens)
        idx = np.array([j for l in lists for j in l], dtype=np.int32)
        self.links[which] = (off, idx)

    def get_links(self, which: int):
        off, idx = self.links[which]
        return [idx[off[i]:off[i + 1]].tolist() for i in range(self.count)]

    # ------------------------------------------------------------------ defaults
    def defaults(self) -> "AgentState":
        """PhysiCell defaults of a freshly created, default-phenotype cell."""
        n, T = self.count, self.defs
        r = self.real
        ff = r(0.75)
        total = r(2494)
        fluid = ff * total
        nuclear = r(540.0)
        nuclear_fluid = ff * nuclear
        cyto = total - nuclear
        cyto_fluid = ff * cyto
        vals = {
            "radius": 8.412710547954228,
            "nuclear_radius": 5.051670902881889,
            "surface_area": 889.3685284131693,
            "volume": total,
            "fluid_fraction": ff,
            "fluid": fluid,
            "solid": total - fluid,
            "nuclear": nuclear,
            "nuclear_fluid": nuclear_fluid,
            "nuclear_solid": nuclear - nuclear_fluid,
            "cytoplasmic": cyto,
            "cytoplasmic_fluid": cyto_fluid,
            "cytoplasmic_solid": cyto - cyto_fluid,
            "target_solid_cytoplasmic": cyto - cyto_fluid,
            "target_solid_nuclear": nuclear - nuclear_fluid,
            "cytoplasmic_to_nuclear_ratio": cyto / (r(1e-16) + nuclear),
            "cell_cell_adhesion_strength": 0.4,
            "cell_cell_repulsion_strength": 10.0,
            "cell_BM_repulsion_strength": 100.0,
            "relative_maximum_adhesion_distance": 1.25,
            "attachment_elastic_constant": 0.01,
            "attachment_rate": 0.0,
            "detachment_rate": 0.0,
            "persistence_time": 1.0,
            "migration_speed": 1.0,
            "migration_bias": 0.0,
            "damage_rate": 1.0,
            "dead_phagocytosis_rate": 0.0,
            "number_of_nuclei": 1,
            "is_movable": 1,
            "maximum_number_of_attachments": 12,
            "chemotaxis_direction": 1,
        }
        for k, v in vals.items():
            self.arrays[k][:] = v
        self.arrays["cell_adhesion_affinities"][:] = 1
        self.arrays["immunogenicities"][:] = 1
        del n, T
        return self


def pack_lists(lists, count):
    lens = np.array([len(l) for l in lists], dtype=np.int64)
    off = np.zeros(count + 1, np.int32)
    off[1:] = np.cumsum(lens)
    idx = np.array([j for l in lists for j in l], dtype=np.int32)
    return off, idx


def unpack_lists(off, idx):
    return [idx[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]
