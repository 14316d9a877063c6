"""ORACLE (test infrastructure only): Python wrapper of liboracle.so.

Same interface as paraphysicell_amd.mechanics.DeviceSim, so a test can run the device
library and the CPU restatement side by side on identical host arrays.  Only tests/,
bench.py's cpu_baseline leg and __graft_entry__.smoke() import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

from paraphysicell_amd import abi
from paraphysicell_amd.mechanics import _LibSim, _declare

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

_LIB = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    args = ["make", "-C", ORACLE_DIR]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return ORACLE_LIB


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        lib = ctypes.CDLL(ORACLE_LIB)
        _declare(lib, "oracle_")
        lib.oracle_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.oracle_set_rng.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64]
        lib.oracle_time_steps.argtypes = [ctypes.c_void_p, ctypes.POINTER(abi.StepParams), ctypes.c_int32,
                                          ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
        lib.oracle_max_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


class OracleSim(_LibSim):
    """CPU restatement of the reference mechanics step, single-threaded by default."""

    prefix = "oracle_"
    # the oracle's mesh/neighbor phases take an (ignored) params pointer like the others
    _mesh_only_phases = ()

    def __init__(self, mesh: abi.MeshSpec, real_bytes: int = 4, defs: int = 1, subs: int = 1,
                 capacity: int = 1024, seed: int = 0, threads: int = 1, rng: str = "keyed"):
        lib = load()
        super().__init__(lib, mesh, real_bytes, defs, subs, capacity, seed)
        lib.oracle_set_threads(self._h, int(threads))
        self.set_rng(rng, seed)

    def set_rng(self, mode: str, seed: int):
        self._lib.oracle_set_rng(self._h, 1 if mode == "mt" else 0, int(seed))

    def set_threads(self, threads: int):
        self._lib.oracle_set_threads(self._h, int(threads))

    def time_steps(self, params, warmup: int, steps: int) -> float:
        ms = ctypes.c_double()
        p = params.to_c()
        self._lib.oracle_time_steps(self._h, ctypes.byref(p), int(warmup), int(steps), ctypes.byref(ms))
        return ms.value

    @staticmethod
    def max_threads() -> int:
        return load().oracle_max_threads()
