"""Host check of the scan kernels' cell walk: bonxai_amd/csrc/bxg_walk.h,
the same source the GPU library compiles, built with the host compiler and
run over random rays (leaf sizes 2^1..2^4, short and long rays, walks from
the origin and from a seek, with and without the endpoint cell) against the
closed form of RayIterator's cells (probabilistic_map.hpp:162-203).  The
closed form is checked equal to the oracle's ray in test_oracle.py."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cell_runs_match_closed_form():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "ray_walk_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "bonxai_amd", "csrc"),
                        os.path.join(HERE, "walk", "ray_walk_check.cpp"), "-o", exe], check=True)
        for seed in (1, 2, 3):
            out = subprocess.run([exe, str(seed), "20000"], capture_output=True, text=True, timeout=120)
            assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
