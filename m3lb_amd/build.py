"""Builds libm3lb_hip.so for gfx950 with hipcc.

Fourteen instantiation units (one per lattice x collision operator, from
csrc/inst.hip), the C-ABI unit of the single-box solver (csrc/solver.hip) and
the batched server (csrc/serve.hip) compile in parallel into build objects under
m3lb_amd/_build/, then link into m3lb_amd/libm3lb_hip.so.  Every HIP line
carries -Xclang -target-feature -Xclang -packed-fp32-ops: hipcc then emits no
packed FP32 instructions.
"""
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libm3lb_hip.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument"]

UNITS = [(q, op) for q in (9, 27) for op in range(7)]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return cmd


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC)
            if f.endswith((".hpp", ".h"))] + [os.path.join(ROOT, "include", "m3lb.h")]


def build_id():
    """Hash of every device source, the public header and the compile flags:
    the identity of the step kernels a build contains.  Baked into the library
    (m3lb_kernel_build_id) so that a measurement can be matched to the kernels
    it was taken on."""
    h = hashlib.sha256()
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hpp", ".h", ".hip")):
            h.update(f.encode())
            with open(os.path.join(CSRC, f), "rb") as fh:
                h.update(fh.read())
    with open(os.path.join(ROOT, "include", "m3lb.h"), "rb") as fh:
        h.update(fh.read())
    h.update(" ".join(FLAGS[:-3]).encode())
    return h.hexdigest()[:16]


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(force=True, jobs=None, verbose=False):
    """Compiles every unit (force=True) or only those older than a source."""
    os.makedirs(OBJ, exist_ok=True)
    headers = _headers()
    jobs = jobs or min(16, os.cpu_count() or 4)
    cmds = []
    objs = []
    for q, op in UNITS:
        o = os.path.join(OBJ, f"inst_q{q}_op{op}.o")
        objs.append(o)
        if force or _stale(o, headers + [os.path.join(CSRC, "inst.hip")]):
            cmds.append([HIPCC] + FLAGS + [f"-DM3LB_Q={q}", f"-DM3LB_OP={op}", "-c",
                                           os.path.join(CSRC, "inst.hip"), "-o", o])
    every = headers + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip")]
    for unit in ("solver", "serve"):
        o = os.path.join(OBJ, f"{unit}.o")
        objs.append(o)
        # the C-ABI unit carries the build id, so it is rebuilt when any source changes
        deps = every if unit == "solver" else headers + [os.path.join(CSRC, f"{unit}.hip")]
        if force or _stale(o, deps):
            extra = [f'-DM3LB_BUILD_ID="{build_id()}"'] if unit == "solver" else []
            cmds.append([HIPCC] + FLAGS + extra + ["-c", os.path.join(CSRC, f"{unit}.hip"), "-o", o])
    # largest units first
    cmds.sort(key=lambda c: ("q27" not in c[-1], c[-1]))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for c in ex.map(_run, cmds):
            if verbose:
                print("built", c[-1], flush=True)
    if force or cmds or not os.path.exists(LIB):
        tmp = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print(LIB)
