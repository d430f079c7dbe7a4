"""bench.py's command line on the CPU: the options parse, and a request for
more than one GPU exits with a message before touching a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_help_lists_the_legs():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--serving-small-requests", "--no-e2m13", "--e2m13-steps", "--serving-no-graph",
                "--precision-study", "--configs"):
        assert opt in r.stdout


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
