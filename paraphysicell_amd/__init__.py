"""MI355X-native mechanics timestep of paraPhysiCell (src/solver/), behind a C-ABI.

Host side: ``mechanics.DeviceSim`` / ``mechanics.Solver`` over libpcm.so,
``state.AgentState`` for the reference's SoA layout, ``synth`` for the synthetic
populations of the benchmark configurations.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
