"""The CPU oracle against the reference's own committed results.

tests/golden/reference_data.json holds data the reference checked in (written
by tests/golden/make_golden.py): the grid-convergence ladders, the 2D
Taylor-Green files, the Galilean-invariance table and the Poiseuille tables.
The oracle must reproduce each to the digits the reference printed.
"""
import math

import pytest

from golden_cases import (GAL, GOLD, OPS, PRINT_TOL, TG2, decaying, poiseuille, rel,
                          tgv2d_run)
from m3lb_amd import abi, cases
from oracle.binding import OracleSolver


@pytest.mark.parametrize("op", list(OPS))
def test_taylor_green_ladder_d2q9(op):
    """doc/fig/conv_taylor_green_<op>.dat (validation/grid_convergence.cpp:39-47)."""
    for N, err in GOLD["grid_convergence"]["taylor_green"][op]["rows"]:
        got, _, _ = decaying("taylor_green", abi.D2Q9, OPS[op], int(N), solver_cls=OracleSolver, threads=8)
        assert rel(got, err) < PRINT_TOL, (op, N, got, err)


@pytest.mark.parametrize("op", list(OPS))
def test_taylor_green_d3q27_lies_on_d2q9(op):
    """doc/m3lb.tex:3884-3886: the z-independent field gives the same error on
    D3Q27 (Nz = 8) as on D2Q9."""
    N, err = GOLD["grid_convergence"]["taylor_green"][op]["rows"][0]
    got, _, _ = decaying("taylor_green", abi.D3Q27, OPS[op], int(N), solver_cls=OracleSolver, threads=8)
    assert rel(got, err) < PRINT_TOL


@pytest.mark.parametrize("op", list(OPS))
def test_abc_ladder_d3q27(op):
    """doc/fig/conv_abc_<op>.dat (validation/grid_convergence.cpp:50-57)."""
    for N, err in GOLD["grid_convergence"]["abc"][op]["rows"]:
        if N > 32:
            continue  # N = 48 runs in the slow set
        got, _, _ = decaying("abc", abi.D3Q27, OPS[op], int(N), solver_cls=OracleSolver, threads=8)
        assert rel(got, err) < PRINT_TOL, (op, N, got, err)


@pytest.mark.slow
@pytest.mark.parametrize("op", list(OPS))
def test_abc_ladder_d3q27_n48(op):
    N, err = GOLD["grid_convergence"]["abc"][op]["rows"][-1]
    got, _, _ = decaying("abc", abi.D3Q27, OPS[op], int(N), solver_cls=OracleSolver, threads=8)
    assert rel(got, err) < PRINT_TOL




@pytest.mark.parametrize("op", list(OPS))
def test_galilean_nu_eff(op):
    """README.md Galilean table (validation/galilean.cpp:41-69): nu_eff of the
    ABC flow under a uniform translation U0, six printed digits."""
    g = GOLD["galilean_nu_eff"]
    for U0, want in zip(g["U0"], g["rows"][GAL[op]]):
        _, nu_eff, _ = decaying("abc", abi.D3Q27, OPS[op], 32, U=0.01, U0=U0, solver_cls=OracleSolver, threads=8)
        assert abs(nu_eff - want) <= 5.01e-7, (op, U0, nu_eff, want)




@pytest.mark.parametrize("name", list(TG2))
def test_tgv2d_results(name):
    """results/A_taylor_green/tgv2d_<name>.dat; "bgk" is the campaign's
    product-form BGK (validation/Campaign.hpp:36)."""
    lat, op = TG2[name]
    for N, _tau, steps, err in GOLD["tgv2d"][name]["rows"]:
        if N > 16:
            continue
        got, T = tgv2d_run(lat, op, int(N), solver_cls=OracleSolver, threads=8)
        assert T == int(steps)
        assert rel(got, err) < PRINT_TOL, (name, N, got, err)


@pytest.mark.parametrize("col", ["BGK", "TRT", "MRT", "CM"])
def test_poiseuille_slip_table(col):
    """README.md slip table (validation/poiseuille.cpp:203-218): the printed
    slip (3 digits) where it is a discretisation defect, |slip| < 1e-11 (the
    reference's FP64 acceptance, poiseuille.cpp:195) where it is round-off."""
    g = GOLD["poiseuille_slip"]
    ci = g["columns"].index(col)
    for row in g["rows"]:
        tau, want = row[0], row[ci]
        if abs(tau - 0.9330127) < 1e-7:
            tau = 0.5 + math.sqrt(3.0) / 4.0
        _, rho_mean, slip, _ = poiseuille(OPS[col], tau, solver_cls=OracleSolver)
        if abs(want) < 1e-11:
            assert abs(slip) < 1e-11, (col, tau, slip)
        else:
            assert abs(slip - want) <= 0.0051 * abs(want), (col, tau, slip, want)
        assert abs(rho_mean - 1.0) < 1e-9  # poiseuille.cpp:321


def test_poiseuille_bgk_l2_table():
    """README.md BGK L2 table (poiseuille.cpp:221-238), H = 32 rows."""
    for tau, H, l2, slip in GOLD["poiseuille_bgk_l2"]["rows"]:
        if int(H) != 32:
            continue
        if abs(tau - 0.9330127) < 1e-7:
            tau = 0.5 + math.sqrt(3.0) / 4.0
        got, _, s, _ = poiseuille(abi.OP_BGK, tau, solver_cls=OracleSolver)
        if l2 < 1e-11:
            assert got < 1e-11
        else:
            assert abs(got - l2) <= 0.0051 * l2
            assert abs(s - slip) <= 0.0051 * abs(slip)


def test_poiseuille_streaming_and_storage_equivalence():
    """validation/poiseuille.cpp:305-316: TwoLattice and EsotericPull give the
    same L2 bit for bit, raw and shifted."""
    for storage in (abi.RAW, abi.SHIFTED):
        a = poiseuille(abi.OP_CM, 0.875, storage=storage, solver_cls=OracleSolver, two_lattice=True)
        b = poiseuille(abi.OP_CM, 0.875, storage=storage, solver_cls=OracleSolver, two_lattice=False)
        assert a[0] == b[0]
        assert a[0] < 1e-11


@pytest.mark.slow
def test_tgv3d_first_probe():
    """results/E_tgv3d/tgv3d_re1600_d3q27_cm.dat (validation/tgv3d.cpp:55-104):
    D = 64, CM, tau = 0.5024; E/E0 and enstrophy ratio at the first probe
    (t* = 0.5, 1600 steps), nine printed digits.  The GPU tests run the file's
    remaining probes."""
    D, u0, Re = 64, 0.02, 1600.0
    nu = u0 * D / Re
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, (D, D, D),
                          omega=1.0 / (nu * 3.0 + 0.5))
    s = OracleSolver(cfg, threads=8)
    s.initialize_field(*cases.tgv3d_campaign(cfg, u0))
    e0, w0, ok0 = s.energy_enstrophy()
    s.step(1600)
    e1, w1, ok1 = s.energy_enstrophy()
    t, E, W = GOLD["tgv3d_re1600_d3q27_cm"]["rows"][1]
    assert t == 0.5 and ok0 and ok1
    assert rel(e1 / e0, E) < PRINT_TOL
    assert rel(w1 / w0, W) < PRINT_TOL


def _forcing_rows(H_max):
    return [r for r in GOLD["forcing_transverse"]["rows"] if r[1] <= H_max]


@pytest.mark.parametrize("row", _forcing_rows(16), ids=lambda r: f"tau{r[0]}-H{int(r[1])}")
def test_forcing_transverse_golden(row):
    """results/H_forcing_convergence/forcing_conv.dat (forcing_transverse.cpp
    -conv): both orientations of the forced D2Q9 CM channel; at tau = 7/8 both
    are exact to round-off."""
    from golden_cases import forcing_transverse
    tau, H = row[0], int(row[1])
    for orient, (l2_ref, slip_ref) in ((0, (row[2], row[4])), (1, (row[5], row[7]))):
        l2, slip = forcing_transverse(OracleSolver, orient, tau, H)
        if abs(tau - 0.875) < 1e-12:
            assert l2 < 1e-11 and l2_ref < 1e-11, (orient, l2, l2_ref)
            assert abs(slip) < 1e-11, (orient, slip)
        else:
            assert rel(l2, l2_ref) < PRINT_TOL, (orient, l2, l2_ref)
            assert rel(slip, slip_ref) < 1e-6, (orient, slip, slip_ref)


@pytest.mark.parametrize("tau", [0.875, 1.2])
def test_forcing_3d_channels_every_axis(tau):
    """forcing_3d.cpp PART 1 (:230-255): the three cyclic wall/force
    orientations agree to round-off, and at tau = 7/8 each is exact."""
    from golden_cases import forcing_3d_channel
    l2 = [forcing_3d_channel(OracleSolver, c, (c + 1) % 3, tau, H=8) for c in range(3)]
    assert max(l2) - min(l2) <= 1e-12, l2
    if abs(tau - 0.875) < 1e-12:
        assert max(l2) < 1e-11, l2
    else:
        assert min(l2) > 1e-6  # away from the magic point the parabola is not exact


def test_forcing_3d_sine_converges_second_order():
    """forcing_3d.cpp PART 2 (:256-270) on the oracle: L2 error of the exact
    3D forced solution falls at second order, N = 8 -> 16 -> 32."""
    from golden_cases import forcing_3d_sine
    l2 = [forcing_3d_sine(OracleSolver, N, 0.8, threads=8)[0] for N in (8, 16, 32)]
    orders = [math.log(a / b) / math.log(2.0) for a, b in zip(l2, l2[1:])]
    # 1.90 from 8 to 16 (pre-asymptotic), 1.98 from 16 to 32
    assert 1.85 < orders[0] < 2.1 and 1.95 < orders[1] < 2.05, (l2, orders)
    assert l2[-1] < 1e-2


def test_constant_force_field_is_uniform_guo():
    """A per-node force that is the same everywhere is uniform Guo, bit for bit."""
    import numpy as np
    shape = (6, 5, 4)
    F = (2e-5, -1e-5, 3e-6)
    a = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, shape, omega=1.3, force=F)
    b = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, shape, omega=1.3,
                        force="field")
    fa, fb = OracleSolver(a), OracleSolver(b)
    n = fa.dom[9]
    fb.set_force_field(*(np.full(n, v) for v in F))
    for s in (fa, fb):
        s.initialize_field(*cases.taylor_green_box(a, 0.03))
        s.step(7)
    assert np.array_equal(fa.gather_populations(), fb.gather_populations())
    assert fa.total_momentum(0) == fb.total_momentum(0)


def test_shear_layer_first_probes():
    """results/B_shear_layer/shear_d2q9_cm.dat (shear_layer.cpp): energy and
    enstrophy ratios of the first four probes (1280 steps) to the printed digits."""
    from golden_cases import shear_layer_run
    rows = GOLD["shear_layer"]["d2q9_cm"]["rows"]
    got = shear_layer_run(OracleSolver, abi.D2Q9, 4, threads=8)
    for (t, e, w), ref in zip(got, rows):
        assert abs(t - ref[0]) < 1e-9
        assert rel(e, ref[1]) < PRINT_TOL, (t, e, ref)
        assert rel(w, ref[2]) < PRINT_TOL, (t, w, ref)
