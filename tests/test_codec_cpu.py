"""Host check of the narrow 16-bit float (E2M13 storage, and the same codec
with 3 exponent bits): every code against its field definition, re-encoding,
nearest-even rounding of random floats and at every tie, saturation.  The
device kernels include the same header."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_narrow_float_codec(tmp_path):
    exe = tmp_path / "narrow_float_test"
    src = os.path.join(ROOT, "tests", "cpp", "narrow_float_test.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", src, "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
