"""The order in which the one-launch step deals tiles to blocks (and so to the
8 XCDs) changes only which block updates a node, never the update itself:
the slot image after several steps must be bit-identical for every tile
order, including the defaults the solver picks per storage type (runs of 64
tiles for FP32, 8 for E2M13, 2 for FP16), on rows of one and of two 256-node
chunks and with a Solid obstacle."""
import sys

import numpy as np
import pytest

from m3lb_amd import abi, cases
from m3lb_amd.solver import FluidSolver

pytestmark = pytest.mark.gpu

ORDERS = [0, -1, 1, 2, 3, 8, 64, 128]


@pytest.fixture(autouse=True)
def announce(request):
    print(f"\n[gpu-case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def image(cfg, run, flags, steps=5):
    with FluidSolver(cfg) as s:
        s.set_step_fusion(abi.FUSION_OFF)
        if run is not None:
            s.set_tile_order(run)
        if flags is not None:
            s.set_geometry(flags)
        s.initialize_field(*cases.tgv3d_benchmark(cfg, 0.04, perturbation=1e-3, seed=7))
        s.step(steps)
        return s.storage()


@pytest.mark.parametrize("dtype", [abi.F32, abi.E2M13, abi.F16])
@pytest.mark.parametrize("shape", [(64, 20, 12), (300, 14, 9)])
def test_every_tile_order_gives_the_same_bits(dtype, shape):
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, shape, omega=1.7)
    fl = np.zeros(shape[::-1], dtype=np.uint8)
    fl[3:6, 4:9, 10:20] = abi.SOLID
    ref = image(cfg, None, fl)  # the default order
    assert np.isfinite(ref).all()
    for run in ORDERS:
        assert np.array_equal(image(cfg, run, fl), ref), f"tile order {run}"
