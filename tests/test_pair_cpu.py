"""CPU check of the two-step pass's schedule (m3lb_amd/csrc/pair_plan.hpp)
through a g++-built program: ticket coverage, waits on earlier tickets only,
race-free slot access within a launch and one stripe per slot, and the same
array as two whole steps when the tiles run in ticket order
(tests/cpp/pair_plan_test.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pair_schedule(tmp_path):
    exe = tmp_path / "pair_plan_test"
    src = os.path.join(ROOT, "tests", "cpp", "pair_plan_test.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", src, "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
