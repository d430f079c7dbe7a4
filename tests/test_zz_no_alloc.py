"""m3lb_totals_static, the entry point that allocates nothing, against the
entry points it stands in for (m3lb_total_mass and m3lb_total_momentum, the
reference's reduce<P>, FluidSolver.hpp:601-606, 1205-1229): the same bits,
on two calls in a row, the second on a smaller box, so that the static
scratch array holds the first call's partials when the second one starts."""
import sys

import numpy as np
import pytest

from m3lb_amd import abi
from m3lb_amd.solver import FluidSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    print(f"\n[gpu-case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def box(shape, dtype):
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, shape, periodic=(1, 0, 1),
                          omega=1.5, force=(1e-5, 2e-6, 0.0))
    nx, ny, nz = shape
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, 0, :] = fl[:, ny - 1, :] = abi.SOLID
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    fl[(x - nx // 3) ** 2 + (y - ny // 2) ** 2 + (z - nz // 2) ** 2 <= 16] = abi.SOLID
    return cfg, fl


@pytest.mark.parametrize("dtype", [abi.F32, abi.F64])
def test_totals_static_matches_the_reductions(dtype):
    for shape in ((200, 96, 80), (24, 18, 10)):  # 4096 reduction blocks, then 17
        cfg, fl = box(shape, dtype)
        with FluidSolver(cfg) as s:
            s.set_geometry(fl)
            s.initialize_taylor_green(0.04)
            s.step(7)
            got = s.totals_static()
            want = np.array([s.total_mass()] + [s.total_momentum(a) for a in range(3)])
            assert np.array_equal(got, want), (shape, got, want)
            assert got[0] > 0.9 * shape[0] * shape[1] * shape[2] * 0.9
