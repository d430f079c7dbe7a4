"""m3lb_momentum_exchange on the GPU against the numpy restatement of
validation/square_cylinder.cpp:340-350 (over the oracle's and over the
device's own gathered populations), and the steady-channel force balance."""
import numpy as np
import pytest

from golden_cases import momentum_exchange_np
from m3lb_amd import abi, cases
from m3lb_amd.solver import FluidSolver
from oracle.binding import OracleSolver

pytestmark = pytest.mark.gpu

TOL = {abi.F64: 1e-12, abi.F32: 2e-5}


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("lat,op,storage", [(abi.D3Q27, abi.OP_CM, abi.SHIFTED),
                                            (abi.D3Q27, abi.OP_BGK, abi.RAW),
                                            (abi.D2Q9, abi.OP_TRT, abi.SHIFTED)])
def test_drag_matches_restatement(lat, op, storage, dtype):
    shape = (40, 30, 24) if lat == abi.D3Q27 else (80, 40, 1)
    cfg = abi.make_config(lat, op, storage, dtype, shape, periodic=(1, 0, 0 if shape[2] > 1 else 1),
                          omega=1.5, omega_minus=1.1, force=(2e-5, 0.0, 0.0))
    if lat == abi.D3Q27:
        flags, body, _ = cases.sphere_channel(shape, n_spheres=3, rmin=3, rmax=5, seed=2)
    else:
        fl = cases.channel_flags(shape).reshape(40, 80)
        yy, xx = np.mgrid[0:40, 0:80]
        body = (((xx - 20) ** 2 + (yy - 19) ** 2) <= 36).astype(np.uint8)
        fl[body != 0] = abi.SOLID
        flags, body = fl.reshape(-1), body.reshape(-1)
    with FluidSolver(cfg) as s:
        o = OracleSolver(cfg)
        for x in (s, o):
            x.set_geometry(flags)
            x.initialize(1.0, 0.01, 0.0, 0.0)
            x.step(25)
        g_dev = s.gather_populations()
        g_orc = o.gather_populations()
        for b in (body, None):
            F = np.array(s.momentum_exchange(b))
            bb = (flags == abi.SOLID) if b is None else b
            own, mag = momentum_exchange_np(cfg, flags, bb, g_dev, magnitude=True)
            orc = momentum_exchange_np(cfg, flags, bb, g_orc)
            # error scale: the summed magnitude of the link terms, which
            # cancel to a small F for raw (unshifted) populations
            scale = np.max(mag)
            assert scale > 0
            assert np.max(np.abs(F - own)) <= 1e-13 * scale, (F, own)
            assert np.max(np.abs(F - orc)) <= TOL[dtype] * scale, (F, orc)


@pytest.mark.parametrize("lat,dtype,tol", [(abi.D2Q9, abi.F64, 1e-9), (abi.D3Q27, abi.F64, 1e-9),
                                           (abi.D3Q27, abi.F32, 1e-3)])
def test_steady_channel_force_balance(lat, dtype, tol):
    """square_cylinder.cpp:44-50: at steady state the walls carry G * N_fluid."""
    G = 1e-6
    shape = (8, 18, 1) if lat == abi.D2Q9 else (8, 18, 8)
    cfg = abi.make_config(lat, abi.OP_CM, abi.SHIFTED, dtype, shape, periodic=(1, 0, 1),
                          omega=1.0, force=(G, 0.0, 0.0))
    flags = cases.channel_flags(shape)
    with FluidSolver(cfg) as s:
        s.set_geometry(flags)
        s.initialize(1.0)
        s.step(12000)  # ~ 20 diffusion times (H = 16, nu = 1/6)
        F = s.momentum_exchange()
    target = G * int(np.sum(flags == abi.FLUID))
    assert abs(F[0] - target) <= tol * target, (F, target)
    assert abs(F[1]) <= tol * target and abs(F[2]) <= tol * target
