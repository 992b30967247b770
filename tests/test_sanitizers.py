"""CPU builds under AddressSanitizer and UndefinedBehaviorSanitizer (gcc
-fsanitize=address,undefined, no recovery, leak checking on):
- the oracle (oracle/psp_oracle.c) through its whole path, tests/sanitize_oracle.c;
- the host planning code (psp_amd/csrc/symbolic.cpp: panel split, nested
  dissection, symbolic analysis, front schedules, solve tiles and the fused
  work lists) through tests/fused_plan_check.cpp, which also checks every
  index the solve kernels read against its array;
- the grid container I/O (psp_amd/csrc/grid_io.cpp), tests/sanitize_grid_io.cpp.
An invalid access, a leak or an undefined operation (signed overflow, bad
shift, misaligned load, out-of-range conversion) fails the run.  The device
code has no sanitizer on this hardware; its indexing is covered by the same
plan checks on the host."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")


def _build(out, cmd):
    r = subprocess.run(cmd + ["-o", out], capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip("the compiler has no sanitizer runtime: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    return out


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "psp_amd", "csrc")]
    oracle = _build(str(d / "sanitize_oracle"),
                    ["gcc", "-std=gnu11", "-fcx-fortran-rules", "-fopenmp"] + SAN +
                    ["-I", os.path.join(ROOT, "oracle"),
                     os.path.join(ROOT, "oracle", "psp_oracle.c"),
                     os.path.join(ROOT, "tests", "sanitize_oracle.c"), "-lm"])
    plan = _build(str(d / "fused_plan_check"),
                  ["g++", "-std=c++17"] + SAN + inc +
                  [os.path.join(ROOT, "tests", "fused_plan_check.cpp"),
                   os.path.join(ROOT, "psp_amd", "csrc", "symbolic.cpp")])
    gio = _build(str(d / "sanitize_grid_io"),
                 ["g++", "-std=c++17"] + SAN + inc +
                 [os.path.join(ROOT, "tests", "sanitize_grid_io.cpp"),
                  os.path.join(ROOT, "psp_amd", "csrc", "grid_io.cpp")])
    return dict(oracle=oracle, plan=plan, gio=gio, dir=str(d))


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    return r.stdout


@pytest.mark.parametrize("grid", ["12 10 24 3", "9 13 27 1", "16 16 20 8", "10 14 20 2"])
def test_oracle_under_sanitizers(exes, grid):
    out = _run([exes["oracle"]] + grid.split())
    assert "iterations" in out


@pytest.mark.parametrize("args", ["40 40 40", "9 13 27", "100 100 100", "40 40 24 12 32 4",
                                  "24 22 30 8 16", "30 20 26 1"])
def test_host_planning_under_sanitizers(exes, args):
    out = _run([exes["plan"]] + args.split())
    assert "class 0: cut" in out


def test_grid_io_under_sanitizers(exes):
    assert "ok" in _run([exes["gio"], exes["dir"]])
