"""Host check of the device RNG engines (hdrestir_amd/csrc/mt64.h, compiled for the host with hipcc)
against libstdc++ std::mt19937_64: every output of the first-twist engine with spill slots and of
the deep-path engine (lazy second twist, spill slot from the third), random draw / discard
patterns and replay copies, over 1400 outputs per seed (tests/rng_deep_check.cpp)."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_device_rng_engines_match_mt19937_64():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "rng_deep_check")
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "rng_deep_check.cpp")],
                       check=True, capture_output=True, cwd=d)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
