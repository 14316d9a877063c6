"""The counter-based stream shared by the device kernels and the oracle's keyed mode."""
import ctypes

import numpy as np

from oracle import pyoracle


def philox(ctr, key):
    lib = pyoracle.load()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    out = (ctypes.c_uint32 * 4)()
    lib.oracle_philox(c, k, out)
    return [int(x) for x in out]


def test_philox4x32_10_known_answers(oracle_lib):
    """Known-answer vectors published with the Random123 reference implementation."""
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_uniformity(oracle_lib):
    """Output words behave like independent uniform 32-bit draws (chi-square on 16 bins)."""
    words = []
    for a in range(2000):
        words += philox([a, 7, 3, 1 << 16], [11, 13])
    w = np.array(words, dtype=np.uint64)
    hist = np.bincount((w >> 28).astype(np.int64), minlength=16)
    expected = len(w) / 16
    chi2 = ((hist - expected) ** 2 / expected).sum()
    assert chi2 < 40  # 15 dof, p ~ 5e-4
