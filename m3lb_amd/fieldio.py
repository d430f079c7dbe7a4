"""Field output in the reference's file formats, so fields from this build
open in the same tools as the reference's and compare line for line.

  write_dat   the campaign results table: '# ' header line, then one row of
              space-separated numbers per line (validation/Campaign.hpp:133-139,
              e.g. results/E_tgv3d/tgv3d_re1600_d3q27_cm.dat)
  read_dat    the rows of such a file (header and comment lines skipped)
  write_vti   VTK ImageData, ASCII Float64 rho, u (3 components) and UInt8
              flag over the interior, x fastest (src/io/VtiWriter.hpp:14-60)
  write_vti_bin  VTK ImageData with appended raw Float32 arrays, each
              preceded by its UInt32 byte count (src/io/VtiWriter.hpp:62-117)
  write_pvd   a time series collection of .vti files (src/io/VtiWriter.hpp:119-134)
  read_vti    the arrays of a file written by either writer

Host-side only: the arrays come from FluidSolver.compute_macroscopic() (padded
index) and are cut to the interior here.
"""
import re
import struct

import numpy as np

from . import abi


def write_dat(path, header, rows, fmt="%.8e"):
    """header: column names (list) or the full header text without '# '.
    fmt: one printf format for every column, or one per column (the
    campaign's tables print t* as %.6f and the ratios as %.8e,
    validation/tgv3d.cpp:79-92)."""
    if not isinstance(header, str):
        header = " ".join(header)
    with open(path, "w") as f:
        f.write("# " + header + "\n")
        for r in rows:
            fm = [fmt] * len(r) if isinstance(fmt, str) else list(fmt)
            if len(fm) != len(r):
                raise ValueError(f"write_dat: {len(fm)} formats for a row of {len(r)} values")
            f.write(" ".join(c % float(v) for c, v in zip(fm, r)) + "\n")


def read_dat(path):
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or line[0].isalpha():
                continue
            out.append([float(v) for v in line.split()])
    return out


def interior(cfg, a):
    """Interior values of a padded per-node array, x fastest."""
    nx, ny, nz, hx, hy, hz, sx, sy, sz, _ = abi.domain_of(cfg)
    return np.asarray(a).reshape(sz, sy, sx)[hz:hz + nz, hy:hy + ny, hx:hx + nx].reshape(-1)


def _extent(nx, ny, nz):
    return f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"


def write_vti(path, cfg, rho, ux, uy, uz, flags):
    """ASCII ImageData of the interior: rho, u and the cell flag.  rho, u*
    are padded arrays (n_padded); flags are interior flags (x fastest)."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    r = interior(cfg, rho)
    u = np.stack([interior(cfg, ux), interior(cfg, uy), interior(cfg, uz)], axis=1)
    fl = np.asarray(flags, dtype=np.uint8).reshape(-1)
    g = lambda v: "%.9g" % v
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n')
        f.write('<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">\n')
        f.write(f'  <ImageData WholeExtent="{_extent(nx, ny, nz)}" Origin="0 0 0" '
                f'Spacing="1 1 1">\n')
        f.write(f'    <Piece Extent="{_extent(nx, ny, nz)}">\n'
                '      <PointData Scalars="rho" Vectors="u">\n')
        f.write('        <DataArray type="Float64" Name="rho" format="ascii">\n')
        f.write("".join(g(v) + " " for v in r))
        f.write("\n        </DataArray>\n")
        f.write('        <DataArray type="Float64" Name="u" NumberOfComponents="3" '
                'format="ascii">\n')
        f.write("".join(g(a) + " " + g(b) + " " + g(c) + " " for a, b, c in u))
        f.write("\n        </DataArray>\n")
        f.write('        <DataArray type="UInt8" Name="flag" format="ascii">\n')
        f.write("".join(str(int(v)) + " " for v in fl))
        f.write("\n        </DataArray>\n"
                "      </PointData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n")


def write_vti_bin(path, shape, arrays, spacing=1.0):
    """arrays: list of (name, data) with data of shape (n,) or (n, 3), n =
    nx*ny*nz in x-fastest order; stored as Float32."""
    nx, ny, nz = shape
    n = nx * ny * nz
    prepared = []
    for name, data in arrays:
        a = np.ascontiguousarray(data, dtype=np.float32)
        ncomp = 1 if a.ndim == 1 else a.shape[1]
        if a.size != n * ncomp:
            raise ValueError(f"write_vti_bin: array '{name}' has {a.size} values, "
                             f"expected {n * ncomp}")
        prepared.append((name, ncomp, a.reshape(-1)))
    scal = next((nm for nm, c, _ in prepared if c == 1), "")
    vec = next((nm for nm, c, _ in prepared if c == 3), "")
    sp = "%g" % spacing
    with open(path, "wb") as f:
        w = lambda s: f.write(s.encode())
        w('<?xml version="1.0"?>\n')
        w('<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" '
          'header_type="UInt32">\n')
        w(f'  <ImageData WholeExtent="{_extent(nx, ny, nz)}" Origin="0 0 0" '
          f'Spacing="{sp} {sp} {sp}">\n')
        w(f'    <Piece Extent="{_extent(nx, ny, nz)}">\n      <PointData')
        if scal:
            w(f' Scalars="{scal}"')
        if vec:
            w(f' Vectors="{vec}"')
        w(">\n")
        off = 0
        for name, ncomp, a in prepared:
            w(f'        <DataArray type="Float32" Name="{name}" NumberOfComponents="{ncomp}" '
              f'format="appended" offset="{off}"/>\n')
            off += 4 + a.size * 4
        w('      </PointData>\n    </Piece>\n  </ImageData>\n'
          '  <AppendedData encoding="raw">\n_')
        for _, _, a in prepared:
            f.write(struct.pack("<I", a.size * 4))
            f.write(a.astype("<f4").tobytes())
        w("\n  </AppendedData>\n</VTKFile>\n")


def write_pvd(path, frames):
    """frames: list of (time, file name)."""
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n'
                '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
                "  <Collection>\n")
        for t, name in frames:
            f.write(f'    <DataSet timestep="{"%.9g" % t}" group="" part="0" file="{name}"/>\n')
        f.write("  </Collection>\n</VTKFile>\n")


def read_vti(path):
    """Returns (extent (nx, ny, nz), {name: array}) of a file written by
    write_vti or write_vti_bin (vector arrays as (n, 3))."""
    raw = open(path, "rb").read()
    head_end = raw.find(b"<AppendedData")
    text = raw[:head_end if head_end >= 0 else len(raw)].decode()
    m = re.search(r'WholeExtent="0 (\d+) 0 (\d+) 0 (\d+)"', text)
    nx, ny, nz = (int(v) + 1 for v in m.groups())
    out = {}
    if head_end >= 0:
        base = raw.index(b"_", head_end) + 1
        for dm in re.finditer(r'<DataArray type="Float32" Name="([^"]+)" NumberOfComponents="(\d+)"'
                              r' format="appended" offset="(\d+)"/>', text):
            name, ncomp, off = dm.group(1), int(dm.group(2)), int(dm.group(3))
            nb = struct.unpack("<I", raw[base + off:base + off + 4])[0]
            a = np.frombuffer(raw[base + off + 4:base + off + 4 + nb], dtype="<f4").copy()
            out[name] = a.reshape(-1, ncomp) if ncomp > 1 else a
    else:
        for dm in re.finditer(r'<DataArray type="(\w+)" Name="([^"]+)"(?: NumberOfComponents="(\d+)")?'
                              r' format="ascii">\n(.*?)\n\s*</DataArray>', text, re.S):
            typ, name, ncomp = dm.group(1), dm.group(2), int(dm.group(3) or 1)
            dt = np.uint8 if typ == "UInt8" else np.float64
            a = np.array(dm.group(4).split(), dtype=np.float64).astype(dt)
            out[name] = a.reshape(-1, ncomp) if ncomp > 1 else a
    return (nx, ny, nz), out
