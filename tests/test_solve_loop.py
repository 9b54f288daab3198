"""CPU test of the pressure solves' host control loops (hnanosolver_amd/csrc/hns_solve_loop.hpp):
builds tests/cpp/solve_loop_test.cpp with g++ and runs it. The program drives the library's own loop
templates with simulated devices: normal convergence, recurrence residuals drifting below the true
one, breakdowns that repeat at one iteration, a device counter that never moves, NaN residuals and
right-hand sides, and the iteration cap. Every case must end within its launch budget."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_solve_loops_end_on_breakdown_stuck_counter_and_nan(tmp_path):
    exe = str(tmp_path / "solve_loop_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", os.path.join(ROOT, "tests", "cpp",
                                                                              "solve_loop_test.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "solve loops: ok" in r.stdout
