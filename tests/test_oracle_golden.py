"""The CPU oracle against the reference's own committed results.

tests/golden/reference_data.json holds data the reference checked in (written
by tests/golden/make_golden.py): the grid-convergence ladders, the 2D
Taylor-Green files, the Galilean-invariance table and the Poiseuille tables.
The oracle must reproduce each to the digits the reference printed.
"""
import json
import math
import os

import numpy as np
import pytest

from m3lb_amd import abi, cases
from oracle.binding import OracleSolver

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_data.json")))
OPS = {"BGK": abi.OP_BGK, "TRT": abi.OP_TRT, "MRT": abi.OP_MRT, "CM": abi.OP_CM}


def magic_minus(w):
    return 1.0 / ((3.0 / 16.0) / (1.0 / w - 0.5) + 0.5)


def decaying(case, lat, op, N, tau=0.8, frac=0.5, U=None, U0=0.0, dtype=abi.F64,
             solver_cls=OracleSolver, **kw):
    """validation/DecayingFlows.hpp:128-198 + run_and_measure (:93-119).
    Returns (l2_u, nu_eff, steps)."""
    nu = (tau - 0.5) * (1.0 / 3.0)
    k = 2.0 * math.pi / N
    coef = 2.0 * k * k if case == "taylor_green" else k * k
    steps = cases.decaying_steps(nu, coef, frac)
    w = 1.0 / tau
    shape = (N, N, (8 if lat == abi.D3Q27 else 1)) if case == "taylor_green" else (N, N, N)
    cfg = abi.make_config(lat, op, abi.SHIFTED, dtype, shape, omega=w,
                          omega_minus=magic_minus(w))
    s = solver_cls(cfg, **kw)
    if case == "taylor_green":
        U = 0.64 / N if U is None else U
        s.initialize_field(*cases.taylor_green(cfg, U))
    else:
        U = 0.32 / N if U is None else U
        s.initialize_field(*cases.abc_flow(cfg, U, U0))
    idx = cases.interior_index(cfg)

    def norm():
        _, ux, uy, uz = s.compute_macroscopic()
        return math.sqrt(float(np.sum((ux[idx] - U0) ** 2 + (uy[idx] - U0) ** 2 +
                                      (uz[idx] - U0) ** 2)))

    n0 = norm()
    s.step(steps)
    n1 = norm()
    nu_eff = -math.log(n1 / n0) / (coef * steps)
    _, ux, uy, uz = s.compute_macroscopic()
    E = math.exp(-nu * coef * steps)
    ana = (cases.tg_analytic(cfg, U, k, E) if case == "taylor_green"
           else cases.abc_analytic(cfg, U, k, E))
    if U0:
        ana = tuple(a + U0 for a in ana)
    return cases.relative_l2(cfg, ux, uy, uz, ana), nu_eff, steps


def rel(a, b):
    return abs(a - b) / abs(b)


# 9 printed significant digits: agreement to 1e-8 relative is the printing limit.
PRINT_TOL = 2e-8


@pytest.mark.parametrize("op", list(OPS))
def test_taylor_green_ladder_d2q9(op):
    """doc/fig/conv_taylor_green_<op>.dat (validation/grid_convergence.cpp:39-47)."""
    for N, err in GOLD["grid_convergence"]["taylor_green"][op]["rows"]:
        got, _, _ = decaying("taylor_green", abi.D2Q9, OPS[op], int(N), threads=8)
        assert rel(got, err) < PRINT_TOL, (op, N, got, err)


@pytest.mark.parametrize("op", list(OPS))
def test_taylor_green_d3q27_lies_on_d2q9(op):
    """doc/m3lb.tex:3884-3886: the z-independent field gives the same error on
    D3Q27 (Nz = 8) as on D2Q9."""
    N, err = GOLD["grid_convergence"]["taylor_green"][op]["rows"][0]
    got, _, _ = decaying("taylor_green", abi.D3Q27, OPS[op], int(N), threads=8)
    assert rel(got, err) < PRINT_TOL


@pytest.mark.parametrize("op", list(OPS))
def test_abc_ladder_d3q27(op):
    """doc/fig/conv_abc_<op>.dat (validation/grid_convergence.cpp:50-57)."""
    for N, err in GOLD["grid_convergence"]["abc"][op]["rows"]:
        if N > 32:
            continue  # N = 48 runs in the slow set
        got, _, _ = decaying("abc", abi.D3Q27, OPS[op], int(N), threads=8)
        assert rel(got, err) < PRINT_TOL, (op, N, got, err)


@pytest.mark.slow
@pytest.mark.parametrize("op", list(OPS))
def test_abc_ladder_d3q27_n48(op):
    N, err = GOLD["grid_convergence"]["abc"][op]["rows"][-1]
    got, _, _ = decaying("abc", abi.D3Q27, OPS[op], int(N), threads=8)
    assert rel(got, err) < PRINT_TOL


GAL = {"BGK": "BGK", "TRT": "TRT(3/16)", "MRT": "MRT (raw)", "CM": "CentralMoments"}


@pytest.mark.parametrize("op", list(OPS))
def test_galilean_nu_eff(op):
    """README.md Galilean table (validation/galilean.cpp:41-69): nu_eff of the
    ABC flow under a uniform translation U0, six printed digits."""
    g = GOLD["galilean_nu_eff"]
    for U0, want in zip(g["U0"], g["rows"][GAL[op]]):
        _, nu_eff, _ = decaying("abc", abi.D3Q27, OPS[op], 32, U=0.01, U0=U0, threads=8)
        assert abs(nu_eff - want) <= 5.01e-7, (op, U0, nu_eff, want)


def tgv2d_run(lat, op, N, Re=1000.0, u0=0.02, dtype=abi.F64, solver_cls=OracleSolver, **kw):
    """validation/tgv2d.cpp:69-157: error at t = T against exp(-1) decay."""
    xi = 2.0 * math.pi / N
    nu = u0 * N / Re
    T = int(1.0 / (2.0 * xi * xi * nu))
    cfg = abi.make_config(lat, op, abi.SHIFTED, dtype, (N, N, 1), omega=1.0 / (nu * 3.0 + 0.5))
    s = solver_cls(cfg, **kw)
    s.initialize_field(*cases.tgv2d(cfg, u0))
    s.step(T)
    _, ux, uy, _ = s.compute_macroscopic()
    x, y, _ = cases.interior_xyz(cfg)
    dec = math.exp(-1.0)
    ax = -u0 * np.cos(xi * x) * np.sin(xi * y) * dec
    ay = u0 * np.sin(xi * x) * np.cos(xi * y) * dec
    idx = cases.interior_index(cfg)
    num = float(np.sum((ux[idx] - ax) ** 2 + (uy[idx] - ay) ** 2))
    return math.sqrt(num / float(np.sum(ax * ax + ay * ay))), T


TG2 = {"d2q9_cm": (abi.D2Q9, abi.OP_CM), "d3q27_bgk": (abi.D3Q27, abi.OP_BGK_PRODUCT),
       "d3q27_cm": (abi.D3Q27, abi.OP_CM), "d3q27_mrt": (abi.D3Q27, abi.OP_MRT)}


@pytest.mark.parametrize("name", list(TG2))
def test_tgv2d_results(name):
    """results/A_taylor_green/tgv2d_<name>.dat; "bgk" is the campaign's
    product-form BGK (validation/Campaign.hpp:36)."""
    lat, op = TG2[name]
    for N, _tau, steps, err in GOLD["tgv2d"][name]["rows"]:
        if N > 16:
            continue
        got, T = tgv2d_run(lat, op, int(N), threads=8)
        assert T == int(steps)
        assert rel(got, err) < PRINT_TOL, (name, N, got, err)


def poiseuille(op, tau, H=32, dtype=abi.F64, storage=abi.SHIFTED, solver_cls=OracleSolver,
               **kw):
    cfg, fl, nu, G = cases.poiseuille_config(op, tau, H, dtype=dtype, storage=storage)
    s = solver_cls(cfg, **kw)
    s.set_geometry(fl)
    s.initialize(1.0)
    return cases.poiseuille_run(s, cfg, H, nu, G)


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
        _, rho_mean, slip, _ = poiseuille(OPS[col], tau)
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
        got, _, s, _ = poiseuille(abi.OP_BGK, tau)
        if l2 < 1e-11:
            assert got < 1e-11
        else:
            assert abs(got - l2) <= 0.0051 * l2
            assert abs(s - slip) <= 0.0051 * abs(slip)


def test_poiseuille_streaming_and_storage_equivalence():
    """validation/poiseuille.cpp:305-316: TwoLattice and EsotericPull give the
    same L2 bit for bit, raw and shifted."""
    for storage in (abi.RAW, abi.SHIFTED):
        a = poiseuille(abi.OP_CM, 0.875, storage=storage, two_lattice=True)
        b = poiseuille(abi.OP_CM, 0.875, storage=storage, two_lattice=False)
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
