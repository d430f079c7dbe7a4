"""Device fields through the reference's file formats: the campaign results
table (validation/Campaign.hpp:133-139, validation/tgv3d.cpp:79-92) and VTK
ImageData (src/io/VtiWriter.hpp), written from what the C-ABI returns and read
back equal to the oracle's fields and to the reference's own results file."""
import math
import sys

import numpy as np
import pytest

from golden_cases import GOLD, PRINT_TOL, rel, tgv3d_run
from m3lb_amd import abi, cases, fieldio
from m3lb_amd.solver import FluidSolver
from oracle.binding import OracleSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    print(f"\n[gpu-case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def _channel():
    shape = (24, 20, 16)
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, shape, periodic=(1, 0, 1),
                          omega=1.3, force=(1e-5, 0.0, 0.0))
    fl = cases.channel_flags(shape).reshape(shape[::-1])
    z, y, x = np.meshgrid(*(np.arange(n) for n in shape[::-1]), indexing="ij")
    fl[(x - 8) ** 2 + (y - 9) ** 2 + (z - 7) ** 2 <= 9] = abi.SOLID
    return cfg, fl


def test_device_fields_through_vti_match_oracle(tmp_path):
    """A forced channel with a sphere, 30 steps: the device's macroscopic
    fields written as ASCII and as appended-raw .vti read back equal to the
    oracle's fields to the files' own precision (9 significant digits;
    Float32), the flags exactly, and a .pvd names both."""
    cfg, fl = _channel()
    field = cases.taylor_green_box(cfg, 0.03)
    o = OracleSolver(cfg)
    o.set_geometry(fl)
    o.initialize_field(*field)
    o.step(30)
    want = o.compute_macroscopic()
    with FluidSolver(cfg) as s:
        s.set_geometry(fl)
        s.initialize_field(*field)
        s.step(30)
        rho, ux, uy, uz = s.compute_macroscopic()
    a = tmp_path / "channel_0030.vti"
    fieldio.write_vti(a, cfg, rho, ux, uy, uz, fl)
    ext, arr = fieldio.read_vti(a)
    assert ext == (cfg.nx, cfg.ny, cfg.nz)
    fluid = fl.reshape(-1) == abi.FLUID
    np.testing.assert_allclose(arr["rho"], fieldio.interior(cfg, want[0]), rtol=1e-8, atol=0)
    for k in range(3):
        w = fieldio.interior(cfg, want[1 + k])
        np.testing.assert_allclose(arr["u"][fluid, k], w[fluid], rtol=1e-8, atol=1e-17)
    assert np.array_equal(arr["flag"], fl.reshape(-1))
    b = tmp_path / "channel_0030_bin.vti"
    u = np.stack([fieldio.interior(cfg, v) for v in (ux, uy, uz)], axis=1)
    fieldio.write_vti_bin(b, ext, [("rho", fieldio.interior(cfg, rho)), ("u", u)])
    _, arr2 = fieldio.read_vti(b)
    np.testing.assert_allclose(arr2["rho"], fieldio.interior(cfg, want[0]).astype(np.float32),
                               rtol=2.4e-7, atol=0)
    wu = np.stack([fieldio.interior(cfg, v) for v in want[1:]], axis=1).astype(np.float32)
    np.testing.assert_allclose(arr2["u"], wu, rtol=2.4e-7, atol=1e-12)
    p = tmp_path / "channel.pvd"
    fieldio.write_pvd(p, [(30.0, a.name), (30.0, b.name)])
    assert f'file="{b.name}"' in open(p).read()


def test_device_results_table_matches_reference_file(tmp_path):
    """validation/tgv3d.cpp's results table written from the device's run of
    the same case (D = 64, Re = 1600, u0 = 0.02, CM FP64), first three probes:
    the header and row layout of results/E_tgv3d/tgv3d_re1600_d3q27_cm.dat,
    and its values to the printed digits."""
    g = GOLD["tgv3d_re1600_d3q27_cm"]
    got = tgv3d_run(FluidSolver, 2)
    D, Re, u0 = 64, 1600.0, 0.02
    tau = 3.0 * u0 * D / Re + 0.5
    header = f"t*  E/E0  Psi/Psi0   D={D} Re={Re:.0f} u0={u0:.3f} tau={tau:.6f}"
    p = tmp_path / "tgv3d_re1600_d3q27_cm.dat"
    fieldio.write_dat(p, header, got, fmt=("%.6f", "%.8e", "%.8e"))
    lines = open(p).read().splitlines()
    assert lines[0] == g["header"]
    assert lines[1] == "0.000000 1.00000000e+00 1.00000000e+00"
    for line in lines[1:]:
        t, e, w = line.split()
        assert len(t.split(".")[1]) == 6 and "e" in e and len(e.split("e")[0]) == 10
    back = fieldio.read_dat(p)
    for (t, E, W), (tg, Eg, Wg) in zip(back, g["rows"]):
        assert abs(t - tg) < 1e-9
        assert rel(E, Eg) < PRINT_TOL and rel(W, Wg) < PRINT_TOL, (t, E, Eg, W, Wg)
    assert math.isclose(back[1][0], 0.5)
