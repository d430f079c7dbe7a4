"""The C++ facade (include/m3lb/lbm.hpp) on the device: the Poiseuille program
of tests/cpp/poiseuille_facade.cpp, written as the reference's
validation/poiseuille.cpp drives FluidSolver, reproduces the reference's
README slip table (validation/poiseuille.cpp:203-218); TRT with the
product-form equilibrium (TRT<L, HighOrderEquilibrium<L>, ...>) gives TRT's row."""
import os
import subprocess
import sys

import pytest

from golden_cases import GOLD

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "poiseuille_facade")


def test_facade_poiseuille_matches_reference_table():
    print("\n[gpu-case] facade poiseuille", flush=True)
    sys.stdout.flush()
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in r.stdout.split("\n"):
        if line.strip():
            op, tau, l2, slip, steps = line.split()
            got[(op, round(float(tau), 6))] = (float(l2), float(slip))
    g = GOLD["poiseuille_slip"]
    for row in g["rows"]:
        tau = 0.5 + 3 ** 0.5 / 4 if abs(row[0] - 0.9330127) < 1e-7 else row[0]
        for col in ("BGK", "TRT", "TRT_PRODUCT", "MRT", "CM"):
            key = (col, round(tau, 6))
            if key not in got:
                continue
            want = row[g["columns"].index("TRT" if col == "TRT_PRODUCT" else col)]
            _, slip = got[key]
            if abs(want) < 1e-11:
                assert abs(slip) < 1e-11, (key, slip)
            else:
                assert abs(slip - want) <= 0.0051 * abs(want), (key, slip, want)
    assert len(got) == 15


def test_facade_per_node_force_policy():
    """tests/cpp/sine_force_facade.cpp: a forcing policy written as the
    reference's SinForce (validation/forcing_3d.cpp:64-91) through the facade
    gives the same steady state as the same case driven from Python."""
    from golden_cases import forcing_3d_sine
    from m3lb_amd.solver import FluidSolver
    exe = os.path.join(os.path.dirname(EXE), "sine_force_facade")
    print("\n[gpu-case] facade sine force", flush=True)
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [line.split() for line in r.stdout.split("\n") if line.strip()]
    assert len(rows) == 4
    for N, tau, l2, _ in rows:
        want, _ = forcing_3d_sine(FluidSolver, int(N), float(tau))
        assert abs(float(l2) - want) <= 1e-6 * want, (N, tau, l2, want)


def test_facade_regularized_walls():
    """tests/cpp/cavity_facade.cpp: walls given as the reference's
    cavity_cm.cpp gives them (a lambda -> WallSpec, velocities deduplicated by
    the facade) step exactly as the same cavity set up through ctypes."""
    import numpy as np
    from golden_cases import interior2d
    from m3lb_amd import abi, cases
    from m3lb_amd.solver import FluidSolver
    exe = os.path.join(os.path.dirname(EXE), "cavity_facade")
    print("\n[gpu-case] facade cavity", flush=True)
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    N, Re, steps = 33, 100.0, 3000
    r = subprocess.run([exe, str(N), str(Re), str(steps)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([float(v) for v in r.stdout.split()])
    U = 0.02
    nu = U * (N - 1) / Re
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64, (N, N, 1),
                          periodic=(False, False, True), omega=1.0 / (3.0 * nu + 0.5))
    with FluidSolver(cfg) as s:
        s.set_regularized_walls(*cases.cavity_walls(N, U))
        s.initialize(1.0)
        s.step(steps)
        ux = interior2d(cfg, s.compute_macroscopic()[1])
    np.testing.assert_array_equal(got, ux[:, N // 2] / U)


def test_facade_outlets_against_reference_results():
    """tests/cpp/inlet_facade.cpp: validation/poiseuille_inlet.cpp written
    against the facade (WallSpec with the outlet's density, the outflow
    order) reproduces results/A_poiseuille_inlet/poiseuille_d2q9_cm.dat at
    Ly = 41 and the order-2 figure."""
    from golden_cases import GOLD
    exe = os.path.join(os.path.dirname(EXE), "inlet_facade")
    print("\n[gpu-case] facade inlet Poiseuille", flush=True)
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    row = GOLD["poiseuille_inlet"]["d2q9"]["rows"][0]
    r = subprocess.run([exe, "41"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ly, steps, err = r.stdout.split()
    assert int(ly) == 41 and int(steps) == int(row[2])
    assert abs(float(err) - row[5]) <= 6e-9 * row[5]
    r = subprocess.run([exe, "41", "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert abs(float(r.stdout.split()[2]) - 1.388e-3) <= 5e-7
