"""The solve work lists (psp_amd/csrc/symbolic.cpp: BuildQueues and
BuildSolveTiles) checked on the host by tests/fused_plan_check.cpp, built
here with g++ from the library's own symbolic.cpp: every row block of every
fused node in exactly one queue, slices tiling their row block, every index
in range, every dependency done earlier in the same queue, and, in the
one-launch-per-level lists, the slices of each split row block placed at
positions equal mod 8 (one XCD), for the benchmark grids, small and ragged
grids, every tile height, a deeper cut, more queues than XCDs, a single
queue, and one-workgroup-per-tile shapes with up to 16 slices; and the
factorization's Schur-complement workspace (updates of even and odd depths
in two regions, each level's updates disjoint, extend-add offsets)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fused") / "fused_plan_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "psp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "fused_plan_check.cpp"),
                    os.path.join(ROOT, "psp_amd", "csrc", "symbolic.cpp"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("args", [
    "100 100 100", "40 40 40", "12 10 24", "9 13 27", "16 16 20", "24 22 30", "64 64 32",
    "200 200 100", "48 40 26", "100 100 100 8 64", "100 100 100 8 16", "40 40 24 12 32 4",
    "40 40 40 8 32 5", "30 20 26 1",
    # one workgroup per tile, slices placed per XCD
    "100 100 100 8 32 -1 64 0 8", "40 40 24 8 32 -1 64 0 16", "9 13 27 8 32 -1 64 0 4",
    "200 200 100 8 32 -1 64 0 8",
])
def test_fused_work_lists(checker, args):
    out = subprocess.run([checker] + args.split(), capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "class 0: cut" in out.stdout and "levels placed per XCD" in out.stdout
    assert "class 0: updates in two regions" in out.stdout
