"""Field output in the reference's formats: round trips through the readers,
the exact header lines of VtiWriter.hpp, and a reference results file read
back through read_dat."""
import os

import numpy as np

from golden_cases import GOLD
from m3lb_amd import abi, cases, fieldio
from oracle.binding import OracleSolver


def test_dat_round_trip(tmp_path):
    p = tmp_path / "tgv3d_d3q27_cm.dat"
    rows = [[0, 1.0, 2.5e-3], [50, 0.99, 2.4e-3]]
    fieldio.write_dat(p, ["t", "E", "Omega"], rows)
    lines = open(p).read().splitlines()
    assert lines[0] == "# t E Omega"
    assert lines[1] == "0.00000000e+00 1.00000000e+00 2.50000000e-03"
    assert fieldio.read_dat(p) == rows


def test_reference_results_file_layout(tmp_path):
    """Rewriting the reference's own tgv3d results rows with its header
    reproduces rows that read back identically."""
    g = GOLD["tgv3d_re1600_d3q27_cm"]
    p = tmp_path / "again.dat"
    fieldio.write_dat(p, g["header"].lstrip("# "), g["rows"])
    assert np.allclose(fieldio.read_dat(p), g["rows"], rtol=1e-8, atol=0)


def _fields():
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, (6, 5, 4),
                          periodic=(1, 0, 1), omega=1.3, force=(1e-5, 0, 0))
    fl = cases.channel_flags((6, 5, 4))
    o = OracleSolver(cfg)
    o.set_geometry(fl)
    o.initialize_field(*cases.taylor_green_box(cfg, 0.03))
    o.step(3)
    return cfg, fl, o.compute_macroscopic()


def test_vti_ascii(tmp_path):
    cfg, fl, (rho, ux, uy, uz) = _fields()
    p = tmp_path / "f.vti"
    fieldio.write_vti(p, cfg, rho, ux, uy, uz, fl)
    text = open(p).read().splitlines()
    assert text[0] == '<?xml version="1.0"?>'
    assert text[1] == '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">'
    assert text[2] == '  <ImageData WholeExtent="0 5 0 4 0 3" Origin="0 0 0" Spacing="1 1 1">'
    ext, arr = fieldio.read_vti(p)
    assert ext == (6, 5, 4)
    assert np.allclose(arr["rho"], fieldio.interior(cfg, rho), rtol=1e-8, atol=0)
    assert np.allclose(arr["u"][:, 0], fieldio.interior(cfg, ux), rtol=1e-8, atol=1e-15)
    assert np.array_equal(arr["flag"], fl)


def test_vti_binary(tmp_path):
    cfg, fl, (rho, ux, uy, uz) = _fields()
    u = np.stack([fieldio.interior(cfg, a) for a in (ux, uy, uz)], axis=1)
    p = tmp_path / "f_bin.vti"
    fieldio.write_vti_bin(p, (6, 5, 4), [("rho", fieldio.interior(cfg, rho)), ("u", u)])
    ext, arr = fieldio.read_vti(p)
    assert ext == (6, 5, 4)
    assert np.array_equal(arr["rho"], fieldio.interior(cfg, rho).astype(np.float32))
    assert np.array_equal(arr["u"], u.astype(np.float32))
    head = open(p, "rb").read(400).decode(errors="replace")
    assert 'header_type="UInt32"' in head and 'Scalars="rho" Vectors="u"' in head
    # the appended block starts with '_' and the first array's byte count
    raw = open(p, "rb").read()
    i = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    assert int.from_bytes(raw[i:i + 4], "little") == 6 * 5 * 4 * 4


def test_pvd(tmp_path):
    p = tmp_path / "series.pvd"
    fieldio.write_pvd(p, [(0.0, "f_0000.vti"), (50.0, "f_0001.vti")])
    t = open(p).read()
    assert '<DataSet timestep="50" group="" part="0" file="f_0001.vti"/>' in t


def test_dat_per_column_formats(tmp_path):
    """The campaign's tables print t* as %.6f and the ratios as %.8e
    (validation/tgv3d.cpp:79-92): the reference's own first rows come back
    character for character."""
    g = GOLD["tgv3d_re1600_d3q27_cm"]
    p = tmp_path / "t.dat"
    fieldio.write_dat(p, g["header"][2:], g["rows"][:2], fmt=("%.6f", "%.8e", "%.8e"))
    lines = open(p).read().splitlines()
    assert lines[0] == g["header"]
    assert lines[1] == "0.000000 1.00000000e+00 1.00000000e+00"
    assert lines[2] == "0.500000 9.04202660e-01 1.61165631e+00"
    import pytest
    with pytest.raises(ValueError):
        fieldio.write_dat(p, "x", [[1.0, 2.0]], fmt=("%.6f",))
