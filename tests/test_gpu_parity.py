"""The HIP path (libm3lb_hip.so, through the C-ABI) against the CPU oracle.

Bit-exact where the work is pure data movement (the esoteric-pull slot map,
seeding, the identity collision); relative L2 of the velocity <= 1e-12 (FP64)
and <= 1e-5 (FP32) where the device computes, the tolerances of the brief's
parity contract.
"""
import itertools
import sys

import numpy as np
import pytest

from m3lb_amd import abi, cases, lattice
from m3lb_amd.solver import FluidSolver
from oracle.binding import OracleSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    print(f"\n[gpu-case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def tag(nx, ny, x, y, z, i, Q):
    return float(((z * ny + y) * nx + x) * Q + i + 1)


PERM_CASES = [(abi.D2Q9, (8, 6, 1), 1), (abi.D2Q9, (8, 6, 1), 2), (abi.D2Q9, (8, 6, 1), 7),
              (abi.D2Q9, (8, 6, 1), 24), (abi.D3Q27, (5, 4, 3), 1), (abi.D3Q27, (5, 4, 3), 13),
              (abi.D3Q27, (5, 4, 3), 2), (abi.D3Q27, (70, 3, 2), 5)]


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("lat,shape,T", PERM_CASES)
def test_permutation(lat, shape, T, dtype):
    """tests/test_streaming.cpp:41-77 on the device: tag every population,
    identity collision, T steps, every value moved exactly T * c_i."""
    nx, ny, nz = shape
    cfg = abi.make_config(lat, abi.OP_NULL, abi.RAW, dtype, shape)
    Q = lattice.Q_of(lat)
    c = lattice.velocities(lat)
    with FluidSolver(cfg) as s:
        seed = np.zeros((s.np, Q))
        for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
            for i in range(Q):
                seed[(z * ny + y) * nx + x, i] = tag(nx, ny, x, y, z, i, Q)
        s.seed_populations(seed)
        s.step(T)
        g = s.gather_populations()
    want = np.zeros_like(g)
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        for i in range(Q):
            want[(z * ny + y) * nx + x, i] = tag(nx, ny, (x - T * c[i, 0]) % nx,
                                                (y - T * c[i, 1]) % ny,
                                                (z - T * c[i, 2]) % nz, i, Q)
    assert (g == want).all()


SLOT_CASES = [(abi.D3Q27, (6, 5, 4), (1, 1, 1)), (abi.D3Q27, (6, 5, 4), (0, 1, 0)),
              (abi.D3Q27, (6, 5, 4), (0, 0, 0)), (abi.D3Q27, (7, 6, 1), (0, 0, 1)),
              (abi.D2Q9, (9, 7, 1), (0, 0, 1)), (abi.D2Q9, (9, 7, 1), (1, 0, 1))]


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("lat,shape,per", SLOT_CASES)
def test_storage_image_bit_exact(lat, shape, per, dtype):
    """The raw slot image after seeding (the reference's ascending scatter with
    halo clamping, FluidSolver.hpp:1231-1241) and after every identity step of
    either parity is bit-identical to the oracle's, halo slots included."""
    cfg = abi.make_config(lat, abi.OP_NULL, abi.RAW, dtype, shape, periodic=per)
    o = OracleSolver(cfg)
    rng = np.random.default_rng(11)
    vals = rng.uniform(-1, 1, size=(o.np, o.Q))
    if dtype == abi.F32:
        vals = vals.astype(np.float32).astype(np.float64)
    with FluidSolver(cfg) as s:
        for phase in range(2):  # seed at even, then again at odd parity
            o.seed_populations(vals)
            s.seed_populations(vals)
            assert (s.storage() == o.storage()).all()
            assert (s.gather_populations() == o.gather_populations()).all()
            for _ in range(3):
                o.step()
                s.step()
                assert (s.storage() == o.storage()).all()
            o.step()
            s.step()
            vals = vals[::-1].copy()


@pytest.mark.parametrize("lat,shape,per", SLOT_CASES[:4])
@pytest.mark.parametrize("op", [abi.OP_BGK, abi.OP_CM])
def test_initialize_matches_oracle(lat, shape, per, op):
    """initialize / initialize_field (FluidSolver.hpp:553-579) with a solid
    block: every slot agrees with the oracle to FP64 rounding."""
    cfg = abi.make_config(lat, op, abi.SHIFTED, abi.F64, shape, periodic=per, omega=1.2)
    o = OracleSolver(cfg)
    fl = np.zeros(shape[::-1], dtype=np.uint8)
    fl[..., 0:2, 1] = abi.SOLID
    o.set_geometry(fl)
    rng = np.random.default_rng(5)
    with FluidSolver(cfg) as s:
        s.set_geometry(fl)
        o.initialize(1.01, 0.02, -0.01, 0.005)
        s.initialize(1.01, 0.02, -0.01, 0.005)
        np.testing.assert_allclose(s.storage(), o.storage(), rtol=0, atol=2e-16)
        field = [1.0 + 0.01 * rng.standard_normal(o.np)] + [0.02 * rng.standard_normal(o.np)
                                                           for _ in range(3)]
        o.initialize_field(*field)
        s.initialize_field(*field)
        np.testing.assert_allclose(s.storage(), o.storage(), rtol=0, atol=2e-16)


def test_identity_mass_drift_is_exactly_zero():
    """Identity collision in FP32: the exact slot sum never changes, and equals
    the oracle's exact sum of the same image."""
    cfg = abi.make_config(abi.D3Q27, abi.OP_NULL, abi.RAW, abi.F32, (40, 24, 16))
    o = OracleSolver(cfg)
    rng = np.random.default_rng(3)
    vals = rng.uniform(0, 0.1, size=(o.np, o.Q)).astype(np.float32).astype(np.float64)
    o.seed_populations(vals)
    with FluidSolver(cfg) as s:
        s.seed_populations(vals)
        before = s.storage_sum_exact()
        assert before == o.storage_sum_exact()
        for _ in range(5):
            s.step(37)
            assert s.storage_sum_exact() == before


def wall_case(lat, dtype, op, storage, force):
    nx, ny, nz = (14, 10, 8) if lat == abi.D3Q27 else (18, 12, 1)
    w = 1.3
    cfg = abi.make_config(lat, op, storage, dtype, (nx, ny, nz), periodic=(1, 0, 1), omega=w,
                          omega_minus=1.0 / ((3.0 / 16.0) / (1.0 / w - 0.5) + 0.5),
                          omega_bulk=1.1 if op in (abi.OP_MRT, abi.OP_CM) else -1.0,
                          force=(2e-5, -1e-5, 3e-6) if force else None)
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, 0, :] = abi.SOLID
    fl[:, ny - 1, :] = abi.SOLID
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    fl[(x - 5) ** 2 + (y - 5) ** 2 + (z - nz // 2) ** 2 <= 4] = abi.SOLID
    return cfg, fl


OPS = [abi.OP_BGK, abi.OP_BGK_PRODUCT, abi.OP_TRT, abi.OP_TRT_PRODUCT, abi.OP_MRT, abi.OP_CM]


@pytest.mark.parametrize("force", [False, True])
@pytest.mark.parametrize("storage", [abi.RAW, abi.SHIFTED])
@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("lat", [abi.D2Q9, abi.D3Q27])
def test_step_matches_oracle(lat, op, dtype, storage, force):
    """run_step with walls, a solid obstacle and Guo forcing, 21 steps (both
    parities): velocity relative L2 against the oracle <= 1e-12 (FP64) /
    1e-5 (FP32); populations and mass likewise."""
    cfg, fl = wall_case(lat, dtype, op, storage, force)
    o = OracleSolver(cfg)
    o.set_geometry(fl)
    rng = np.random.default_rng(1)
    np_ = o.np
    x, y, z, d = cases.padded_coords(cfg)
    field = [1.0 + 0.02 * np.sin(0.7 * x + 0.3 * z),
             0.04 * np.cos(0.5 * y) + 0.005 * rng.standard_normal(np_),
             0.02 * np.sin(0.4 * x), 0.01 * np.cos(0.3 * x + 0.2 * y)]
    o.initialize_field(*field)
    with FluidSolver(cfg) as s:
        s.set_geometry(fl)
        s.initialize_field(*field)
        for chunk in (1, 1, 19):
            o.step(chunk)
            s.step(chunk)
            ro, uxo, uyo, uzo = o.compute_macroscopic()
            rs, uxs, uys, uzs = s.compute_macroscopic()
            num = np.sum((uxs - uxo) ** 2 + (uys - uyo) ** 2 + (uzs - uzo) ** 2)
            den = np.sum(uxo ** 2 + uyo ** 2 + uzo ** 2)
            l2 = np.sqrt(num / den)
            tol = 1e-12 if dtype == abi.F64 else 1e-5
            assert l2 <= tol, (chunk, l2)
            assert np.abs(rs - ro).max() <= (1e-13 if dtype == abi.F64 else 1e-6)
        gp = s.gather_populations()
        go = o.gather_populations()
        scale = np.abs(go).max()
        # FP32: the device contracts a*b+c into FMAs, the oracle rounds each op
        assert np.abs(gp - go).max() <= scale * (1e-13 if dtype == abi.F64 else 2e-5)
        mo, ms = o.total_mass(), s.total_mass()
        assert abs(ms - mo) <= abs(mo) * (1e-13 if dtype == abi.F64 else 1e-6)
        for a in range(3):
            assert abs(s.total_momentum(a) - o.total_momentum(a)) <= (1e-13 if dtype == abi.F64 else 1e-6) * max(1.0, abs(mo))


@pytest.mark.parametrize("shape", [(300, 10, 8), (530, 6, 5)], ids=["chunks2", "chunks3"])
@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("op", [abi.OP_BGK, abi.OP_CM])
def test_wide_rows_match_oracle(shape, dtype, op):
    """Rows of more than one 256-node chunk run the step's scalar-row path
    (two chunks: a shift; three: a division): walls in y, an obstacle, Guo
    force, 21 steps against the oracle at the tolerances above."""
    nx, ny, nz = shape
    cfg = abi.make_config(abi.D3Q27, op, abi.SHIFTED, dtype, shape, periodic=(True, False, True),
                          omega=1.3, omega_bulk=1.1 if op == abi.OP_CM else -1.0,
                          force=(2e-5, -1e-5, 3e-6))
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, 0, :] = abi.SOLID
    fl[:, ny - 1, :] = abi.SOLID
    fl[:, ny // 2, nx // 3:nx // 3 + 3] = abi.SOLID
    o = OracleSolver(cfg)
    o.set_geometry(fl)
    x, y, z, d = cases.padded_coords(cfg)
    field = [1.0 + 0.02 * np.sin(0.05 * x + 0.3 * z), 0.04 * np.cos(0.5 * y + 0.01 * x),
             0.02 * np.sin(0.04 * x), 0.01 * np.cos(0.03 * x + 0.2 * y)]
    o.initialize_field(*field)
    with FluidSolver(cfg) as s:
        s.set_geometry(fl)
        s.initialize_field(*field)
        o.step(21)
        s.step(21)
        _, uxo, uyo, uzo = o.compute_macroscopic()
        _, uxs, uys, uzs = s.compute_macroscopic()
        num = np.sum((uxs - uxo) ** 2 + (uys - uyo) ** 2 + (uzs - uzo) ** 2)
        den = np.sum(uxo ** 2 + uyo ** 2 + uzo ** 2)
        assert np.sqrt(num / den) <= (1e-12 if dtype == abi.F64 else 1e-5)
        gp, go = s.gather_populations(), o.gather_populations()
        assert np.abs(gp - go).max() <= np.abs(go).max() * (1e-13 if dtype == abi.F64 else 2e-5)
    o.close()


@pytest.mark.parametrize("dtype", [abi.F16, abi.E2M13])
def test_wide_rows_16bit_match_fp64_storage(dtype):
    """The 16-bit step on rows of two chunks (the path the 512^3 benchmark
    takes) against FP64 storage of the same Taylor-Green box, 100 steps.
    (y and z extents of 32 and 24 keep the flow from decaying to velocities
    at the 16-bit resolution itself.)"""
    shape = (512, 32, 24)
    out = {}
    for dt in (abi.F64, dtype):
        cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dt, shape, omega=1.6)
        with FluidSolver(cfg) as s:
            s.initialize_taylor_green(0.04)
            s.step(100)
            out[dt] = s.compute_macroscopic()[1:]
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(out[dtype], out[abi.F64]))
    den = sum(float(np.sum(b ** 2)) for b in out[abi.F64])
    assert np.sqrt(num / den) <= 1e-3


def test_store_macroscopic_matches_compute():
    """step(store_macroscopic = true) writes the macroscopic state of the
    populations the step read (FluidSolver.hpp:1033-1036), which is what
    compute_macroscopic returns just before that step."""
    cfg, fl = wall_case(abi.D3Q27, abi.F32, abi.OP_CM, abi.SHIFTED, True)
    with FluidSolver(cfg) as s:
        s.set_geometry(fl)
        s.initialize(1.0, 0.02, 0.0, 0.01)
        s.step(5)
        before = s.compute_macroscopic()
        s.step(1, store_macroscopic=True)
        stored = s.macroscopic_stored()
        for a, b in zip(before, stored):
            np.testing.assert_array_equal(a, b)


def test_energy_enstrophy_matches_oracle():
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, (24, 20, 16), omega=1.7)
    o = OracleSolver(cfg)
    o.initialize_field(*cases.tgv3d_campaign(cfg, 0.03))
    with FluidSolver(cfg) as s:
        s.initialize_field(*cases.tgv3d_campaign(cfg, 0.03))
        o.step(9)
        s.step(9)
        eo, wo, oko = o.energy_enstrophy()
        es, ws, oks = s.energy_enstrophy()
    assert oko and oks
    assert abs(es - eo) <= 1e-12 * eo
    assert abs(ws - wo) <= 1e-12 * wo


def test_device_errors_are_reported():
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, (8, 8, 8))
    with FluidSolver(cfg) as s:
        bad = np.full(8 * 8 * 8, 7, dtype=np.uint8)
        with pytest.raises(RuntimeError):
            s.set_geometry(bad)
        with pytest.raises(RuntimeError):
            s.set_force(1e-5, 0, 0)  # created without Guo forcing
        with pytest.raises(RuntimeError):
            s.macroscopic_stored()
        with pytest.raises(RuntimeError):
            s.step(-1)


@pytest.mark.parametrize("shape", [(40, 24, 16), (130, 20, 12)])
def test_cache_policy_is_bit_identical(shape):
    """m3lb_set_cache_policy changes only how the step kernel's local slots
    are cached: every policy leaves the same bits."""
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F32, shape, omega=1.3,
                          force=(1e-5, 0.0, 0.0))
    out = []
    for pol in (abi.CACHE_AUTO, abi.CACHE_STREAM, abi.CACHE_PLAIN):
        with FluidSolver(cfg) as s:
            s.set_cache_policy(pol)
            s.initialize_taylor_green(0.04)
            s.step(13)
            out.append(s.storage())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
