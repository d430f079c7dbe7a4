"""CPU checks of the batched server: planning code (arena, admission order,
exact divider) through a g++-built test program, the ctypes mirror of the
serving structs, footprints and argument validation (no device call)."""
import ctypes
import os
import subprocess

import pytest

from m3lb_amd import abi, serve, solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_planning_code(tmp_path):
    exe = tmp_path / "serve_plan_test"
    src = os.path.join(ROOT, "tests", "cpp", "serve_plan_test.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", src, "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def test_struct_layouts():
    """The serving structs as a C compiler lays them out on x86-64."""
    assert ctypes.sizeof(abi.ServerConfig) == 64
    assert abi.ServerConfig.budget_bytes.offset == 48
    assert abi.ServerConfig.max_resident_nodes.offset == 56
    assert ctypes.sizeof(abi.Request) == 128
    assert abi.Request.omega.offset == 40
    assert abi.Request.force.offset == 64
    assert abi.Request.arrival.offset == 120
    assert ctypes.sizeof(abi.Result) == 72
    assert abi.Result.mass.offset == 40
    assert ctypes.sizeof(abi.ServerStats) == 128
    assert abi.ServerStats.device_seconds.offset == 96
    assert ctypes.sizeof(abi.TuneResult) == 32


def test_struct_layouts_match_the_header(tmp_path):
    """Same offsets from the header itself, through g++."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "m3lb.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\","
        "sizeof(m3lb_server_config), sizeof(m3lb_request), offsetof(m3lb_request, arrival),"
        "sizeof(m3lb_result), sizeof(m3lb_server_stats), offsetof(m3lb_server_stats, device_seconds),"
        "sizeof(m3lb_tune_result));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert [int(v) for v in out] == [64, 128, 120, 72, 128, 96, 32]


def _cfg(**kw):
    return abi.make_server_config(**kw)


@pytest.mark.parametrize("shape,periodic,dtype", [((32, 32, 32), (1, 1, 1), abi.F32),
                                                  ((33, 50, 191), (1, 0, 1), abi.F32),
                                                  ((64, 64, 1), (1, 1, 1), abi.F64),
                                                  ((40, 24, 16), (1, 0, 1), abi.E2M13)])
def test_footprint(shape, periodic, dtype):
    lat = abi.D2Q9 if shape[2] == 1 else abi.D3Q27
    scfg = _cfg(lattice=lat, dtype=dtype)
    req = abi.make_request(shape, 10, periodic=periodic)
    d = abi.domain_of(abi.request_config(scfg, req))
    np_ = d[9]
    stride = (np_ + 63) // 64 * 64
    elem = {abi.F64: 8, abi.F32: 4}.get(dtype, 2)
    rnd = lambda b: (b + 255) // 256 * 256
    assert serve.footprint(scfg, req) == rnd(stride * lat * elem) + rnd(np_)


@pytest.mark.parametrize("bad", [
    dict(op=abi.OP_NULL), dict(dtype=abi.F16, storage=abi.RAW), dict(dtype=abi.E2M13, storage=abi.RAW),
    dict(dtype=9), dict(max_boxes=0),
    dict(steps_per_chunk=0), dict(steps_per_chunk=2000), dict(blocks_per_cu=0),
    dict(tile_shape=7), dict(max_bypass=-1), dict(budget_bytes=0), dict(forcing=5),
    dict(tile_run=-3),
])
def test_invalid_server_configs_are_rejected(bad):
    L = solver.load()
    scfg = _cfg()
    for k, v in bad.items():
        setattr(scfg, k, v)
    h = solver._vp()
    assert L.m3lb_srv_create(ctypes.byref(scfg), ctypes.byref(h)) == abi.EINVAL
    assert L.m3lb_last_error().decode()
    out = ctypes.c_uint64()
    req = abi.make_request((8, 8, 8), 1)
    assert L.m3lb_srv_footprint(ctypes.byref(scfg), ctypes.byref(req), ctypes.byref(out)) == abi.EINVAL


def test_invalid_requests_are_rejected_by_footprint():
    L = solver.load()
    scfg = _cfg(lattice=abi.D2Q9)
    out = ctypes.c_uint64()
    req = abi.make_request((8, 8, 2), 1)  # D2Q9 needs nz == 1
    assert L.m3lb_srv_footprint(ctypes.byref(scfg), ctypes.byref(req), ctypes.byref(out)) == abi.EINVAL
    req = abi.make_request((0, 8, 1), 1)
    assert L.m3lb_srv_footprint(ctypes.byref(scfg), ctypes.byref(req), ctypes.byref(out)) == abi.EINVAL


def test_null_server_arguments():
    L = solver.load()
    assert L.m3lb_srv_create(None, None) == abi.EINVAL
    assert L.m3lb_srv_submit(None, None, None, None) == abi.EINVAL
    assert L.m3lb_srv_progress(None, 0, None) == abi.EINVAL
    assert L.m3lb_srv_destroy(None) == abi.OK


def test_serving_workload_edges():
    """The small-box workload draws every edge from its range, and the
    default range is unchanged (the seeded stream's first request)."""
    from m3lb_amd import cases
    small = cases.serving_workload(400, seed=5, rate=400.0, edges=(16, 48))
    assert all(16 <= e < 48 for _, r in small for e in (r.nx, r.ny, r.nz))
    assert all(b[0] > a[0] for a, b in zip(small, small[1:]))
    first = cases.serving_workload(200)[0][1]
    assert (first.nx, first.ny, first.nz) == (96, 124, 106)
    assert all(32 <= e < 192 for _, r in cases.serving_workload(200) for e in (r.nx, r.ny, r.nz))
