"""CPU-side checks of the drop-in boundary: libm3lb_hip.so loads, exports every
entry point include/m3lb.h declares, and its host-only logic (extents,
configuration validation) matches the reference's rules.  No call here touches
a GPU."""
import ctypes
import os
import re

import pytest

from m3lb_amd import abi, solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "m3lb.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(m3lb_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(solver.LIB_PATH):
        pytest.fail("libm3lb_hip.so is not built (run __graft_entry__.build())")
    return solver.load()


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    assert set(declared_functions()) == set(solver.SIGNATURES)


def test_config_struct_layout():
    """m3lb_config as laid out by a C compiler on x86-64 (include/m3lb.h)."""
    assert ctypes.sizeof(abi.Config) == 96
    assert abi.Config.omega.offset == 40
    assert abi.Config.forcing.offset == 64
    assert abi.Config.force.offset == 72


@pytest.mark.parametrize("shape,periodic", [((64, 64, 64), (1, 1, 1)),
                                            ((8, 34, 1), (1, 0, 1)),
                                            ((512, 256, 256), (1, 0, 0)),
                                            ((5, 4, 3), (0, 0, 0)),
                                            ((7, 1, 1), (0, 0, 0))])
def test_domain_matches_reference_rules(lib, shape, periodic):
    """Domain.hpp:47-56 / 104-108: halo 1 on non-periodic axes longer than 1."""
    lat = abi.D2Q9 if shape[2] == 1 else abi.D3Q27
    cfg = abi.make_config(lat, abi.OP_CM, abi.SHIFTED, abi.F32, shape, periodic=periodic)
    assert solver.domain(cfg) == abi.domain_of(cfg)


def test_domain_matches_oracle():
    from oracle.binding import OracleSolver
    cfg = abi.make_config(abi.D3Q27, abi.OP_BGK, abi.RAW, abi.F64, (6, 5, 4),
                          periodic=(1, 0, 0))
    assert OracleSolver(cfg).dom == solver.domain(cfg)


@pytest.mark.parametrize("bad", [
    dict(lattice=19),
    dict(op=9),
    dict(storage=3),
    dict(dtype=7),
    dict(dtype=abi.F16, storage=abi.RAW),
    dict(dtype=abi.BF16, op=abi.OP_NULL),
    dict(dtype=abi.I16, storage=abi.RAW),
    dict(dtype=abi.E2M13, storage=abi.RAW),
    dict(dtype=6),
    dict(nx=0),
    dict(lattice=abi.D2Q9, nz=4),
    dict(reserved=1),
    dict(forcing=5),
    dict(omega=float("nan")),
])
def test_invalid_configs_are_rejected(lib, bad):
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (8, 8, 8))
    for k, v in bad.items():
        setattr(cfg, k, v)
    out = (ctypes.c_int64 * 10)()
    rc = lib.m3lb_domain(ctypes.byref(cfg), out)
    assert rc == abi.EINVAL
    assert lib.m3lb_last_error().decode()


def test_null_arguments_are_rejected(lib):
    assert lib.m3lb_create(None, None) == abi.EINVAL
    assert lib.m3lb_step(None, 1, 0) == abi.EINVAL
    assert lib.m3lb_destroy(None) == abi.OK


@pytest.mark.parametrize("run", [0, -1, 1, 2, 3, 7, 16, 64, 512])
def test_tile_order_is_a_permutation(lib, run):
    """Every step launch must visit each tile exactly once, for every grid
    size (including grids not divisible by 8 or by 8 * run)."""
    import numpy as np
    for grid in list(range(1, 70)) + [80, 127, 520, 1000, 4096 + 8 * 17 + 3, 524288]:
        m = solver.tile_order_map(grid, run)
        assert np.array_equal(np.sort(m), np.arange(grid, dtype=np.uint32)), (grid, run)
