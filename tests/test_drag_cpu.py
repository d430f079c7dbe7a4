"""Momentum-exchange drag, pinned on the CPU: the numpy restatement of
validation/square_cylinder.cpp:340-350 over the oracle's gathered
populations must reproduce the force balance the reference's own self-test
relies on (square_cylinder.cpp:44-50): in a periodic channel driven by a
uniform body force G, at steady state the walls carry G * N_fluid exactly."""
import numpy as np
import pytest

from golden_cases import momentum_exchange_np
from m3lb_amd import abi, cases
from oracle.binding import OracleSolver


@pytest.mark.parametrize("lat,op,storage,shape", [
    (abi.D2Q9, abi.OP_BGK, abi.SHIFTED, (4, 10, 1)),
    (abi.D2Q9, abi.OP_TRT, abi.RAW, (3, 12, 1)),
    (abi.D3Q27, abi.OP_CM, abi.SHIFTED, (3, 10, 3)),
])
def test_walls_carry_the_body_force(lat, op, storage, shape):
    G = 1e-6
    cfg = abi.make_config(lat, op, storage, abi.F64, shape, periodic=(1, 0, 1), omega=1.0,
                          omega_minus=1.0, force=(G, 0.0, 0.0))
    flags = cases.channel_flags(shape)
    o = OracleSolver(cfg)
    o.set_geometry(flags)
    o.initialize(1.0)
    o.step(6000)  # ~15 diffusion times H^2 / nu at nu = 1/6
    n_fluid = int(np.sum(flags == abi.FLUID))
    F = momentum_exchange_np(cfg, flags, flags == abi.SOLID, o.gather_populations())
    assert abs(F[0] - G * n_fluid) <= 1e-9 * G * n_fluid, (F, G * n_fluid)
    assert abs(F[1]) <= 1e-9 * G * n_fluid
    assert abs(F[2]) <= 1e-9 * G * n_fluid


def test_no_links_no_force():
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, (4, 4, 4), omega=1.2)
    o = OracleSolver(cfg)
    o.initialize(1.0, 0.01, 0.0, 0.0)
    o.step(3)
    body = np.zeros(64, dtype=np.uint8)
    F = momentum_exchange_np(cfg, np.zeros(64, np.uint8), body, o.gather_populations())
    assert np.all(F == 0)


def test_sphere_channel_geometry():
    flags, body, spheres = cases.sphere_channel((96, 64, 48), n_spheres=12, rmin=3, rmax=6, seed=2)
    f = flags.reshape(48, 64, 96)
    b = body.reshape(48, 64, 96).astype(bool)
    assert len(spheres) == 12
    assert np.all(f[:, 0, :] == abi.SOLID) and np.all(f[:, 63, :] == abi.SOLID)
    assert np.all(f[0] == abi.SOLID) and np.all(f[47] == abi.SOLID)
    # spheres keep a free layer to the walls and to each other
    assert not b[:, :2, :].any() and not b[:, 62:, :].any()
    assert not b[:2].any() and not b[46:].any()
    for k, (a, bb, e, r) in enumerate(spheres):
        for (a2, b2, e2, r2) in spheres[k + 1:]:
            dx = min(abs(a - a2), 96 - abs(a - a2))
            assert dx ** 2 + (bb - b2) ** 2 + (e - e2) ** 2 >= (r + r2 + 1.0) ** 2
    # volume close to the sum of 4/3 pi r^3
    vol = sum(4.0 / 3.0 * np.pi * r ** 3 for *_, r in spheres)
    assert abs(b.sum() - vol) < 0.08 * vol
    assert np.all(f[b] == abi.SOLID)
    # same seed, same geometry
    assert np.array_equal(cases.sphere_channel((96, 64, 48), 12, 3, 6, seed=2)[0], flags)
