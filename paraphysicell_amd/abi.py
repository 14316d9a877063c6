"""ctypes mirror of include/pcm.h: structs, status codes and the agent field table.

The field table restates the X-macro list of include/pcm.h (the ids are the list order);
tests check it against ``pcm_field_info`` exported by the built library.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PCM_OK = 0
PCM_ERR_ARGUMENT = 1
PCM_ERR_DEVICE = 2
PCM_ERR_CAPACITY = 3
PCM_ERR_UNSUPPORTED = 4
PCM_ERR_STATE = 5

T_REAL, T_I32, T_U8 = 0, 1, 2
W_ONE, W_DIMS, W_DEFS, W_SUBS = 0, 1, 2, 3

LINKS_SPRINGS = 0
LINKS_ATTACHED = 1

FLAG_NONE = 0
FLAG_TO_REMOVE = 1
FLAG_TO_DIVIDE = 2

# (name, element type, width kind) in pcm.h order; names are the reference's SoA members
FIELDS = [
    ("positions", T_REAL, W_DIMS),
    ("velocities", T_REAL, W_DIMS),
    ("previous_velocities", T_REAL, W_DIMS),
    ("radius", T_REAL, W_ONE),
    ("nuclear_radius", T_REAL, W_ONE),
    ("surface_area", T_REAL, W_ONE),
    ("volume", T_REAL, W_ONE),
    ("fluid", T_REAL, W_ONE),
    ("solid", T_REAL, W_ONE),
    ("nuclear", T_REAL, W_ONE),
    ("nuclear_fluid", T_REAL, W_ONE),
    ("nuclear_solid", T_REAL, W_ONE),
    ("cytoplasmic", T_REAL, W_ONE),
    ("cytoplasmic_fluid", T_REAL, W_ONE),
    ("cytoplasmic_solid", T_REAL, W_ONE),
    ("fluid_fraction", T_REAL, W_ONE),
    ("cytoplasmic_to_nuclear_ratio", T_REAL, W_ONE),
    ("target_solid_cytoplasmic", T_REAL, W_ONE),
    ("target_solid_nuclear", T_REAL, W_ONE),
    ("cell_cell_adhesion_strength", T_REAL, W_ONE),
    ("cell_cell_repulsion_strength", T_REAL, W_ONE),
    ("cell_BM_repulsion_strength", T_REAL, W_ONE),
    ("relative_maximum_adhesion_distance", T_REAL, W_ONE),
    ("cell_adhesion_affinities", T_REAL, W_DEFS),
    ("maximum_number_of_attachments", T_I32, W_ONE),
    ("attachment_elastic_constant", T_REAL, W_ONE),
    ("attachment_rate", T_REAL, W_ONE),
    ("detachment_rate", T_REAL, W_ONE),
    ("is_motile", T_U8, W_ONE),
    ("persistence_time", T_REAL, W_ONE),
    ("migration_speed", T_REAL, W_ONE),
    ("migration_bias_direction", T_REAL, W_DIMS),
    ("migration_bias", T_REAL, W_ONE),
    ("motility_vector", T_REAL, W_DIMS),
    ("restrict_to_2d", T_U8, W_ONE),
    ("dead", T_U8, W_ONE),
    ("dead_phagocytosis_rate", T_REAL, W_ONE),
    ("live_phagocytosis_rates", T_REAL, W_DEFS),
    ("damage_rate", T_REAL, W_ONE),
    ("attack_rates", T_REAL, W_DEFS),
    ("immunogenicities", T_REAL, W_DEFS),
    ("fusion_rates", T_REAL, W_DEFS),
    ("simple_pressure", T_REAL, W_ONE),
    ("number_of_nuclei", T_I32, W_ONE),
    ("damage", T_REAL, W_ONE),
    ("total_attack_time", T_REAL, W_ONE),
    ("cell_definition_index", T_I32, W_ONE),
    ("is_movable", T_U8, W_ONE),
    ("flags", T_U8, W_ONE),
    ("internalized_substrates", T_REAL, W_SUBS),
    ("fraction_transferred_when_ingested", T_REAL, W_SUBS),
    ("origin", T_I32, W_ONE),
    ("contact_function", T_U8, W_ONE),
    ("chemotaxis_index", T_I32, W_ONE),
    ("chemotaxis_direction", T_I32, W_ONE),
    ("chemotactic_sensitivities", T_REAL, W_SUBS),
    ("migration_bias_direction_function", T_U8, W_ONE),
    ("orientation", T_REAL, W_DIMS),
    ("polarity", T_REAL, W_ONE),
]
CONTACT_NONE = 0  # PCM_CONTACT_NONE
CONTACT_STANDARD_ELASTIC = 1  # PCM_CONTACT_STANDARD_ELASTIC
BIAS_NONE = 0  # PCM_BIAS_NONE
BIAS_CHEMOTAXIS = 1  # PCM_BIAS_CHEMOTAXIS
BIAS_ADVANCED_CHEMOTAXIS = 2  # PCM_BIAS_ADVANCED_CHEMOTAXIS
BIAS_ADVANCED_CHEMOTAXIS_NORMALIZED = 3  # PCM_BIAS_ADVANCED_CHEMOTAXIS_NORMALIZED
FIELD_ID = {name: i for i, (name, _, _) in enumerate(FIELDS)}


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32),
        ("bbox_min", ctypes.c_int32 * 3),
        ("bbox_max", ctypes.c_int32 * 3),
        ("voxel", ctypes.c_int32 * 3),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("mesh", MeshDesc),
        ("real_bytes", ctypes.c_int32),
        ("cell_defs", ctypes.c_int32),
        ("substrates", ctypes.c_int32),
        ("capacity", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
    ]


class StepParams(ctypes.Structure):
    _fields_ = [
        ("mechanics_time_step", ctypes.c_double),
        ("virtual_wall_at_domain_edges", ctypes.c_int32),
        ("automated_spring_adhesion", ctypes.c_int32),
    ]


class Tuning(ctypes.Structure):
    """pcm_tuning: launch shape of the lane-parallel kernels (results do not depend on it)."""

    _fields_ = [
        ("velocity_lanes", ctypes.c_int32),
        ("event_lanes", ctypes.c_int32),
        ("neighbor_waves", ctypes.c_int32),
    ]

    def as_tuple(self):
        return (self.velocity_lanes, self.event_lanes, self.neighbor_waves)


@dataclass
class MeshSpec:
    """Microenvironment box (integer microns) and the mechanics voxel shape."""

    dims: int
    bbox_min: tuple
    bbox_max: tuple
    voxel: tuple = (30, 30, 30)

    def to_c(self) -> MeshDesc:
        m = MeshDesc()
        m.dims = self.dims
        for k in range(3):
            m.bbox_min[k] = int(self.bbox_min[k])
            m.bbox_max[k] = int(self.bbox_max[k])
            m.voxel[k] = int(self.voxel[k])
        return m

    def with_voxel(self, voxel) -> "MeshSpec":
        """The same bounding box with another voxel shape (the diffusion mesh)."""
        return MeshSpec(self.dims, self.bbox_min, self.bbox_max, tuple(voxel))

    def grid_shape(self):
        g = []
        for k in range(3):
            if k < self.dims:
                ext = self.bbox_max[k] - self.bbox_min[k]
                g.append(max(1, (ext + self.voxel[k] - 1) // self.voxel[k]))
            else:
                g.append(1)
        return tuple(g)


@dataclass
class Params:
    """Environment scalars read by the mechanics phases (src/environment.h:30-41)."""

    dt: float = 0.1
    virtual_wall: bool = False
    springs: bool = True

    def to_c(self) -> StepParams:
        p = StepParams()
        p.mechanics_time_step = float(self.dt)
        p.virtual_wall_at_domain_edges = int(bool(self.virtual_wall))
        p.automated_spring_adhesion = int(bool(self.springs))
        return p


def field_dtype(ftype: int, real_bytes: int):
    if ftype == T_REAL:
        return np.float32 if real_bytes == 4 else np.float64
    if ftype == T_I32:
        return np.int32
    return np.uint8


def field_width(wkind: int, dims: int, defs: int, subs: int) -> int:
    return {W_ONE: 1, W_DIMS: dims, W_DEFS: defs, W_SUBS: subs}[wkind]


def make_config(mesh: MeshSpec, real_bytes: int, defs: int, subs: int, capacity: int, seed: int) -> Config:
    c = Config()
    c.mesh = mesh.to_c()
    c.real_bytes = real_bytes
    c.cell_defs = defs
    c.substrates = subs
    c.capacity = capacity
    c.seed = seed
    return c


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libpcm.so")
