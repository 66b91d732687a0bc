"""The server's device-region carving (poisson/arena.h) on the CPU: a
randomised take/give run over a host buffer checked for alignment, bounds,
no overlap, exact accounting, full coalescing and best fit
(tests/native/arena_check.cpp, built with g++)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_arena_bookkeeping(tmp_path):
    exe = str(tmp_path / "arena_check")
    try:
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "multigrid-poisson_amd", "poisson"),
                        os.path.join(HERE, "native", "arena_check.cpp"), "-o", exe],
                       check=True, capture_output=True, text=True)
    except FileNotFoundError:
        pytest.skip("g++ not available")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "arena ok" in out.stdout
