"""The HIP path against the reference's own committed results (FP64, the
precision they were produced in), through the C-ABI.  Same drivers as
tests/test_oracle_golden.py; here every ladder runs to its largest size."""
import math
import sys

import pytest

from golden_cases import (GAL, GOLD, OPS, PRINT_TOL, TG2, decaying, poiseuille, rel,
                          tgv2d_run, tgv3d_run)
from m3lb_amd import abi
from m3lb_amd.solver import FluidSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    print(f"\n[gpu-case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


@pytest.mark.parametrize("op", list(OPS))
def test_taylor_green_ladder(op):
    """doc/fig/conv_taylor_green_<op>.dat, on D2Q9 and (Nz = 8) D3Q27."""
    for N, err in GOLD["grid_convergence"]["taylor_green"][op]["rows"]:
        for lat in (abi.D2Q9, abi.D3Q27):
            got, _, _ = decaying("taylor_green", lat, OPS[op], int(N), solver_cls=FluidSolver)
            assert rel(got, err) < PRINT_TOL, (op, lat, N, got, err)


@pytest.mark.parametrize("op", list(OPS))
def test_abc_ladder(op):
    """doc/fig/conv_abc_<op>.dat, N = 16 ... 48."""
    for N, err in GOLD["grid_convergence"]["abc"][op]["rows"]:
        got, _, _ = decaying("abc", abi.D3Q27, OPS[op], int(N), solver_cls=FluidSolver)
        assert rel(got, err) < PRINT_TOL, (op, N, got, err)


@pytest.mark.parametrize("op", list(OPS))
def test_galilean_nu_eff(op):
    g = GOLD["galilean_nu_eff"]
    for U0, want in zip(g["U0"], g["rows"][GAL[op]]):
        _, nu_eff, _ = decaying("abc", abi.D3Q27, OPS[op], 32, U=0.01, U0=U0,
                                solver_cls=FluidSolver)
        assert abs(nu_eff - want) <= 5.01e-7, (op, U0, nu_eff, want)


@pytest.mark.parametrize("name", list(TG2))
def test_tgv2d_results(name):
    """results/A_taylor_green/tgv2d_<name>.dat, every committed N <= 128."""
    lat, op = TG2[name]
    for N, _tau, steps, err in GOLD["tgv2d"][name]["rows"]:
        if N > 128:
            continue
        got, T = tgv2d_run(lat, op, int(N), solver_cls=FluidSolver)
        assert T == int(steps)
        assert rel(got, err) < PRINT_TOL, (name, N, got, err)


@pytest.mark.parametrize("col", ["BGK", "TRT", "MRT", "CM"])
def test_poiseuille_slip_table(col):
    g = GOLD["poiseuille_slip"]
    ci = g["columns"].index(col)
    for row in g["rows"]:
        tau, want = row[0], row[ci]
        if abs(tau - 0.9330127) < 1e-7:
            tau = 0.5 + math.sqrt(3.0) / 4.0
        _, rho_mean, slip, _ = poiseuille(OPS[col], tau, solver_cls=FluidSolver)
        if abs(want) < 1e-11:
            assert abs(slip) < 1e-11, (col, tau, slip)
        else:
            assert abs(slip - want) <= 0.0051 * abs(want), (col, tau, slip, want)
        assert abs(rho_mean - 1.0) < 1e-9


def test_poiseuille_exact_at_magic_points():
    """validation/poiseuille.cpp:247-301: each operator exact (L2 < 1e-11) at
    its magic point for H = 16, 32, 64; off-magic BGK second order."""
    tb = 0.5 + math.sqrt(3.0) / 4.0
    for op, tau in ((abi.OP_BGK, tb), (abi.OP_TRT, 0.6), (abi.OP_MRT, 0.875), (abi.OP_CM, 0.875)):
        for H in (16, 32, 64):
            l2, _, _, _ = poiseuille(op, tau, H=H, solver_cls=FluidSolver)
            assert l2 < 1e-11, (op, tau, H, l2)
    e = [poiseuille(abi.OP_BGK, 0.6, H=H, solver_cls=FluidSolver)[0] for H in (16, 32, 64)]
    assert abs(math.log(e[1] / e[2]) / math.log(2.0) - 2.0) < 0.25


def test_tgv3d_re1600_file():
    """results/E_tgv3d/tgv3d_re1600_d3q27_cm.dat: all 20 probes (32000 steps at
    64^3, FP64), E/E0 and enstrophy ratio."""
    rows = GOLD["tgv3d_re1600_d3q27_cm"]["rows"]
    got = tgv3d_run(FluidSolver, len(rows) - 1)
    worst = 0.0
    for (t, E, W), (tg, Eg, Wg) in zip(got, rows):
        assert abs(t - tg) < 1e-9
        worst = max(worst, rel(E, Eg), rel(W, Wg))
    print(f"tgv3d worst relative difference {worst:.3e}")
    assert worst < 6e-9  # the printed 9 digits


def test_tgv3d_re30000_file():
    """results/E_tgv3d/tgv3d_re30000_d3q27_cm.dat: all 20 probes (32000
    steps at tau = 0.500128, through the transition to turbulence), E/E0 and
    enstrophy ratio to the printed 9 digits."""
    rows = GOLD["tgv3d_re30000_d3q27_cm"]["rows"]
    got = tgv3d_run(FluidSolver, len(rows) - 1, Re=30000.0)
    devs = []
    for (t, E, W), (tg, Eg, Wg) in zip(got, rows):
        assert abs(t - tg) < 1e-9
        devs.append(max(rel(E, Eg), rel(W, Wg)))
    print("tgv3d Re 30000 relative differences per probe: " +
          " ".join(f"{d:.1e}" for d in devs), flush=True)
    assert max(devs) < 6e-9, devs


@pytest.mark.parametrize("row", [r for r in GOLD["forcing_transverse"]["rows"] if r[1] <= 32],
                         ids=lambda r: f"tau{r[0]}-H{int(r[1])}")
def test_forcing_transverse_golden(row):
    """results/H_forcing_convergence/forcing_conv.dat on the device: both
    orientations of the forced D2Q9 CM channel, exact at tau = 7/8."""
    from golden_cases import forcing_transverse
    tau, H = row[0], int(row[1])
    for orient, (l2_ref, slip_ref) in ((0, (row[2], row[4])), (1, (row[5], row[7]))):
        l2, slip = forcing_transverse(FluidSolver, orient, tau, H)
        if abs(tau - 0.875) < 1e-12:
            assert l2 < 1e-11 and l2_ref < 1e-11, (orient, l2, l2_ref)
            assert abs(slip) < 1e-11, (orient, slip)
        else:
            assert rel(l2, l2_ref) < PRINT_TOL, (orient, l2, l2_ref)
            assert rel(slip, slip_ref) < 1e-6, (orient, slip, slip_ref)


@pytest.mark.parametrize("tau", [0.875, 0.6, 1.2])
def test_forcing_3d_channels_every_axis(tau):
    """forcing_3d.cpp PART 1 (:230-255) at H = 16 on the device."""
    from golden_cases import forcing_3d_channel
    l2 = [forcing_3d_channel(FluidSolver, c, (c + 1) % 3, tau, H=16) for c in range(3)]
    assert max(l2) - min(l2) <= 1e-12, l2
    if abs(tau - 0.875) < 1e-12:
        assert max(l2) < 1e-11, l2
    else:
        assert min(l2) > 1e-6


def test_forcing_3d_sine_converges_second_order():
    """forcing_3d.cpp PART 2 on the device, N = 16, 32, 64 at tau = 0.8 and 1.2."""
    from golden_cases import forcing_3d_sine
    for tau in (0.8, 1.2):
        l2 = [forcing_3d_sine(FluidSolver, N, tau)[0] for N in (16, 32, 64)]
        orders = [math.log(a / b) / math.log(2.0) for a, b in zip(l2, l2[1:])]
        assert all(1.95 < o < 2.05 for o in orders), (tau, l2, orders)


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
def test_force_field_matches_oracle(dtype):
    """A random per-node force on a channel with walls: device against oracle."""
    import numpy as np
    from m3lb_amd import cases
    from oracle.binding import OracleSolver
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, (20, 18, 16),
                          periodic=(1, 0, 1), omega=1.4, force="field")
    rng = np.random.Generator(np.random.MT19937(5))
    n = abi.domain_of(cfg)[9]
    F = [rng.uniform(-2e-5, 2e-5, n) for _ in range(3)]
    fl = cases.channel_flags((20, 18, 16))
    with FluidSolver(cfg) as s:
        o = OracleSolver(cfg)
        for x in (s, o):
            x.set_geometry(fl)
            x.set_force_field(*F)
            x.initialize_field(*cases.taylor_green_box(cfg, 0.03))
            x.step(15)
        _, ux, uy, uz = s.compute_macroscopic()
        _, vx, vy, vz = o.compute_macroscopic()
        ma, mb = s.total_momentum(0), o.total_momentum(0)
    num = np.sum((ux - vx) ** 2 + (uy - vy) ** 2 + (uz - vz) ** 2)
    den = np.sum(vx ** 2 + vy ** 2 + vz ** 2)
    tol = 1e-12 if dtype == abi.F64 else 1e-5
    assert (num / den) ** 0.5 <= tol
    assert abs(ma - mb) <= tol * 100 * abs(mb) + 1e-12


def test_constant_force_field_is_uniform_guo():
    import numpy as np
    from m3lb_amd import cases
    F = (2e-5, -1e-5, 3e-6)
    shape = (24, 20, 16)
    a = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, shape, omega=1.3, force=F)
    b = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, shape, omega=1.3,
                        force="field")
    with FluidSolver(a) as sa, FluidSolver(b) as sb:
        sb.set_force_field(*(np.full(sb.np, v) for v in F))
        for s in (sa, sb):
            s.initialize_field(*cases.taylor_green_box(a, 0.03))
            s.step(7)
        assert np.array_equal(sa.gather_populations(), sb.gather_populations())


@pytest.mark.parametrize("name,lattice", [("d2q9_cm", abi.D2Q9), ("d3q27_cm", abi.D3Q27)])
def test_shear_layer_full_run(name, lattice):
    """results/B_shear_layer/shear_{d2q9,d3q27}_cm.dat: all 41 probes to t0."""
    from golden_cases import shear_layer_run
    rows = GOLD["shear_layer"][name]["rows"]
    got = shear_layer_run(FluidSolver, lattice, 40)
    worst = 0.0
    for (t, e, w), ref in zip(got, rows):
        assert abs(t - ref[0]) < 1e-9
        worst = max(worst, rel(e, ref[1]), rel(w, ref[2]))
    print(f"shear layer {name}: worst relative difference {worst:.2e}")
    assert worst < 1e-6
