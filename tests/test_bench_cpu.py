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


def test_cpu_baseline_uses_every_core_of_the_job(monkeypatch):
    """The CPU baseline runs on every core the job may use (its affinity set),
    bounded by the OMP_NUM_THREADS share a host gives a one-GPU job, and names
    the host's CPU count and model."""
    sys.path.insert(0, ROOT)
    import bench
    usable, logical, model = bench.host_cpu()
    assert 1 <= usable <= logical and model
    monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
    assert bench.cpu_threads() == usable
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    assert bench.cpu_threads() == 1
    monkeypatch.setenv("OMP_NUM_THREADS", str(usable + 5))
    assert bench.cpu_threads() == usable
