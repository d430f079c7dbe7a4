"""The two-step pass (m3lb_set_step_fusion, m3lb_amd/csrc/pair.hpp) against one
launch per step, through the C-ABI.

The pass only reorders node updates that the esoteric pull makes independent,
so everything here is bit-exact: the slot image after any number of steps of
either parity, for every collision operator, storage dtype, forcing and
geometry the pass runs, equals the image the single-step kernel leaves.  The
identity collision additionally moves every tagged population by exactly
T * c_i (tests/test_streaming.cpp:41-77) and keeps the exact slot sum.
"""
import itertools
import sys

import numpy as np
import pytest

from m3lb_amd import abi, cases, lattice
from m3lb_amd.solver import FluidSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    print(f"\n[gpu-case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def run(cfg, steps, fusion, flags=None, init=None, chunks=None, **kw):
    with FluidSolver(cfg) as s:
        s.set_step_fusion(fusion, **kw)
        if flags is not None:
            s.set_geometry(flags)
        if init is None:
            s.initialize_taylor_green(0.04)
        else:
            init(s)
        for n in (chunks or [steps]):
            s.step(n)
        return s.storage(), s.fused_passes(), s.timestep()


def sphere_flags(shape, walls):
    nx, ny, nz = shape
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    if walls:
        fl[:, 0, :] = abi.SOLID
        fl[:, ny - 1, :] = abi.SOLID
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    fl[(x - nx // 3) ** 2 + (y - ny // 2) ** 2 + (z - nz // 2) ** 2 <= 9] = abi.SOLID
    return fl


OPS = [abi.OP_BGK, abi.OP_BGK_PRODUCT, abi.OP_TRT, abi.OP_TRT_PRODUCT, abi.OP_MRT, abi.OP_CM]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
def test_pair_is_bit_identical_every_operator(op, dtype):
    """Every operator in FP64 and FP32, Guo force, a solid sphere, 13 steps
    (six passes and one single step): the same bits as one launch per step."""
    shape = (40, 24, 16)
    cfg = abi.make_config(abi.D3Q27, op, abi.SHIFTED, dtype, shape, omega=1.3,
                          omega_minus=1.0 / ((3.0 / 16.0) / (1.0 / 1.3 - 0.5) + 0.5),
                          omega_bulk=1.1 if op in (abi.OP_MRT, abi.OP_CM) else -1.0,
                          force=(2e-5, -1e-5, 3e-6))
    fl = sphere_flags(shape, walls=False)
    a, na, ta = run(cfg, 13, abi.FUSION_OFF, fl)
    b, nb, tb = run(cfg, 13, abi.FUSION_ON, fl, lag=2)
    assert na == 0 and nb == 6 and ta == tb == 13
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [abi.F16, abi.BF16, abi.I16, abi.E2M13])
def test_pair_is_bit_identical_16bit(dtype):
    shape = (36, 27, 12)
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, shape, omega=1.6)
    a, _, _ = run(cfg, 10, abi.FUSION_OFF)
    b, nb, _ = run(cfg, 10, abi.FUSION_ON, lag=3)
    assert nb == 5
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", [abi.F16, abi.E2M13])
def test_pair_is_bit_identical_16bit_wide_rows(dtype):
    """Rows of two 256-node chunks: one launch per step takes its
    scalar-row path there, the pass its own node body; the bits agree."""
    shape = (300, 16, 10)
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, shape, omega=1.6)
    a, _, _ = run(cfg, 8, abi.FUSION_OFF)
    b, nb, _ = run(cfg, 8, abi.FUSION_ON, lag=2)
    assert nb == 4
    assert np.array_equal(a, b)


@pytest.mark.parametrize("shape,periodic", [((70, 30, 12), (1, 0, 0)), ((300, 12, 9), (1, 1, 1)),
                                            ((33, 19, 21), (0, 1, 0))])
@pytest.mark.parametrize("stripes,lag,bpc,group", [(0, 0, 0, 0), (3, 1, 2, 1), (5, 4, 8, 3)])
def test_pair_shapes_and_tuning(shape, periodic, stripes, lag, bpc, group):
    """Halos on non-periodic axes with walls there, rows of two 256-node
    chunks, uneven stripes, every tuning value: bit-identical, both parities
    first (an odd step before the passes shifts them to odd parity)."""
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, shape, periodic=periodic,
                          omega=1.2, force=(1e-5, 0.0, 2e-6))
    nx, ny, nz = shape
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    if not periodic[1]:
        fl[:, 0, :] = fl[:, ny - 1, :] = abi.SOLID
    if not periodic[2]:
        fl[0, :, :] = fl[nz - 1, :, :] = abi.SOLID
    if not periodic[0]:
        fl[:, :, 0] = fl[:, :, nx - 1] = abi.SOLID
    a, _, _ = run(cfg, 12, abi.FUSION_OFF, fl, chunks=[1, 11])
    b, nb, _ = run(cfg, 12, abi.FUSION_ON, fl, chunks=[1, 11], stripes=stripes, lag=lag,
                   blocks_per_cu=bpc, group=group)
    assert nb == 5
    assert np.array_equal(a, b)


def tag(nx, ny, x, y, z, i, Q):
    return float(((z * ny + y) * nx + x) * Q + i + 1)


@pytest.mark.parametrize("T", [2, 7, 24])
def test_pair_permutation(T):
    """tests/test_streaming.cpp:41-77 through the pass: identity collision,
    tags moved exactly T * c_i, and the exact slot sum unchanged."""
    nx, ny, nz = 9, 12, 8
    cfg = abi.make_config(abi.D3Q27, abi.OP_NULL, abi.RAW, abi.F32, (nx, ny, nz))
    Q = 27
    c = lattice.velocities(abi.D3Q27)
    with FluidSolver(cfg) as s:
        s.set_step_fusion(abi.FUSION_ON, stripes=4, lag=2)
        seed = np.zeros((s.np, Q))
        for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
            for i in range(Q):
                seed[(z * ny + y) * nx + x, i] = tag(nx, ny, x, y, z, i, Q)
        s.seed_populations(seed)
        before = s.storage_sum_exact()
        s.step(T)
        assert s.fused_passes() == T // 2
        assert s.storage_sum_exact() == before
        g = s.gather_populations()
    want = np.zeros_like(g)
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        for i in range(Q):
            want[(z * ny + y) * nx + x, i] = tag(nx, ny, (x - T * c[i, 0]) % nx,
                                                (y - T * c[i, 1]) % ny,
                                                (z - T * c[i, 2]) % nz, i, Q)
    assert (g == want).all()


def test_pair_leaves_the_stored_step_to_a_single_launch():
    """step(n, store_macroscopic=True): the last step runs alone and stores the
    macroscopic fields of the populations it read, as without the pass."""
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (40, 24, 16), omega=1.4)
    out = []
    for mode in (abi.FUSION_OFF, abi.FUSION_ON):
        with FluidSolver(cfg) as s:
            s.set_step_fusion(mode, lag=2)
            s.initialize_taylor_green(0.04)
            s.step(6, store_macroscopic=True)
            out.append((s.macroscopic_stored(), s.storage(), s.fused_passes()))
    assert out[0][2] == 0 and out[1][2] == 2
    for a, b in zip(out[0][0], out[1][0]):
        np.testing.assert_array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1])


def test_pair_runs_by_default_above_the_cache_and_matches():
    """AUTO pairs the steps of a 256^3 box (1.8 GB of populations, beyond the
    256 MB last-level cache) and leaves a 64^3 box to one launch per step;
    the 256^3 result is bit-identical to one launch per step."""
    big = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (256, 256, 256), omega=1.9)
    a, na, _ = run(big, 10, abi.FUSION_AUTO)
    b, nb, _ = run(big, 10, abi.FUSION_OFF)
    assert na == 5 and nb == 0
    assert np.array_equal(a, b)
    del a, b
    small = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (64, 64, 64), omega=1.9)
    _, ns, _ = run(small, 10, abi.FUSION_AUTO)
    assert ns == 0


def test_pair_failed_wait_is_reported_and_recoverable():
    """The failure path, forced from the host: every wait of the pass is made
    unsatisfiable (m3lb_test_fail_fused_waits), the pass drains, and the next
    synchronising call fails with M3LB_EDEVICE.  After switching the hook off
    and seeding again the handle steps correctly."""
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (40, 24, 16), omega=1.4)
    ref, _, _ = run(cfg, 4, abi.FUSION_OFF)
    with FluidSolver(cfg) as s:
        s.set_step_fusion(abi.FUSION_ON, lag=2)
        s.initialize_taylor_green(0.04)
        s.test_fail_fused_waits(True)
        s.step(2)
        with pytest.raises(RuntimeError, match="two-step pass"):
            s.total_mass()
        s.test_fail_fused_waits(False)
        s.initialize_taylor_green(0.04)
        s.step(4)
        assert np.array_equal(s.storage(), ref)


def test_fusion_arguments_are_checked():
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (16, 16, 16))
    with FluidSolver(cfg) as s:
        for bad in (dict(mode=3), dict(mode=-1), dict(mode=0, stripes=65),
                    dict(mode=0, blocks_per_cu=9), dict(mode=0, group=65), dict(mode=0, lag=-1)):
            with pytest.raises(RuntimeError):
                s.set_step_fusion(**bad)
