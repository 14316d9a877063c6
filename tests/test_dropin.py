"""The C++ drop-in (include/pcm_dropin.hpp) driven like the reference's gtests, on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "_build", "test_dropin")


def build_dropin_test() -> str:
    """g++ the drop-in test against libpcm.so (host code only; no HIP compiler needed)."""
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++20", "-fopenmp", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp"), "-L", os.path.join(ROOT, "paraphysicell_amd", "_build"),
           "-lpcm", "-Wl,-rpath,$ORIGIN/../../paraphysicell_amd/_build", "-o", BIN]
    subprocess.run(cmd, check=True)
    return BIN


def test_dropin_builds():
    """The drop-in header compiles with g++ against the oracle's environment stand-in."""
    if not os.path.exists(os.path.join(ROOT, "paraphysicell_amd", "_build", "libpcm.so")):
        pytest.skip("libpcm.so not built")
    assert os.path.exists(build_dropin_test())


@pytest.mark.gpu
def test_dropin_matches_oracle():
    if not os.path.exists(BIN):
        build_dropin_test()
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "4"))
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240, env=env)
    print(res.stdout[-4000:], res.stderr[-2000:])
    assert res.returncode == 0, res.stdout[-4000:]
    assert "PASSED" in res.stdout
