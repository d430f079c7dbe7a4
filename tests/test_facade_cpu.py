"""The C++ facade (include/m3lb/lbm.hpp) compiles against the C-ABI with a
plain host compiler, and rejects policy combinations the device path does not
implement at compile time.  Runs on the CPU only (nothing is executed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
LIBDIR = os.path.join(ROOT, "m3lb_amd")


def build_facade_program(src, out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
           os.path.join(CPP, src), "-o", out, "-L", LIBDIR, "-lm3lb_hip",
           f"-Wl,-rpath,{LIBDIR}"]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_facade_program_compiles(tmp_path):
    r = build_facade_program("poiseuille_facade.cpp", str(tmp_path / "pois"))
    assert r.returncode == 0, r.stderr


def test_unsupported_streaming_is_a_compile_error(tmp_path):
    r = build_facade_program("reject_two_lattice.cpp", str(tmp_path / "bad"))
    assert r.returncode != 0
    assert "implements EsotericPull streaming only" in r.stderr
