"""Extracts the reference's own committed results into tests/golden/reference_data.json.

Run where the reference checkout exists (it reads /root/reference as DATA:
.dat result files and the measured tables of its README).  Tests never run
it; they read the JSON it writes.

Sources (all in the reference checkout):
  doc/fig/conv_taylor_green_{BGK,TRT,MRT,CM}.dat   grid_convergence.cpp, D2Q9, tau=0.8
  doc/fig/conv_abc_{BGK,TRT,MRT,CM}.dat            grid_convergence.cpp, D3Q27, tau=0.8
  results/A_taylor_green/tgv2d_*.dat               validation/tgv2d.cpp, Re=1000 u0=0.02
  results/E_tgv3d/tgv3d_re{1600,30000}_d3q27_cm.dat validation/tgv3d.cpp, D=64
  README.md, Galilean table                        validation/galilean.cpp
  README.md, Poiseuille slip and L2 tables         validation/poiseuille.cpp
  results/H_forcing_convergence/forcing_conv.dat   validation/forcing_transverse.cpp -conv,
                                                   D2Q9 CM, both orientations
  results/B_shear_layer/shear_{d2q9,d3q27}_cm.dat  validation/shear_layer.cpp, L=256 Re=3e4
  results/C_cavity/cavity_re*_{d2q9,d3q27}_cm*.dat validation/cavity_cm.cpp, N=129 u_lid=0.02,
                                                   17 Ghia stations and full centrelines
  results/A_poiseuille_inlet/poiseuille_*_cm.dat   validation/poiseuille_inlet.cpp, Ly ladder
  results/D_dipole/dipole_re*_{d2q9,d3q27}_cm.dat  validation/dipole.cpp, D_lb = 512
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_data.json")


def rows(path):
    out = []
    with open(os.path.join(REF, path)) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or line[0].isalpha():
                continue
            out.append([float(v) for v in line.split()])
    return out


def header(path):
    with open(os.path.join(REF, path)) as f:
        return f.readline().strip()


def md_table_after(lines, start_pattern):
    """Rows of the first markdown table whose header line matches."""
    for i, line in enumerate(lines):
        if re.search(start_pattern, line):
            table = []
            for row in lines[i + 2:]:
                if not row.startswith("|"):
                    break
                table.append([c.strip().strip("*") for c in row.strip().strip("|").split("|")])
            return i + 1, table
    raise KeyError(start_pattern)


def main():
    data = {"source": "alexderosis/M3LB reference checkout, extracted by tests/golden/make_golden.py"}
    conv = {}
    for case in ("taylor_green", "abc"):
        conv[case] = {}
        for op in ("BGK", "TRT", "MRT", "CM"):
            p = f"doc/fig/conv_{case}_{op}.dat"
            conv[case][op] = {"file": p, "rows": rows(p)}
    data["grid_convergence"] = conv

    tg2 = {}
    for name in ("d2q9_cm", "d3q27_bgk", "d3q27_cm", "d3q27_mrt"):
        p = f"results/A_taylor_green/tgv2d_{name}.dat"
        tg2[name] = {"file": p, "header": header(p), "rows": rows(p)}
    data["tgv2d"] = tg2

    p = "results/E_tgv3d/tgv3d_re1600_d3q27_cm.dat"
    data["tgv3d_re1600_d3q27_cm"] = {"file": p, "header": header(p), "rows": rows(p)}
    p = "results/E_tgv3d/tgv3d_re30000_d3q27_cm.dat"
    data["tgv3d_re30000_d3q27_cm"] = {"file": p, "header": header(p), "rows": rows(p)}

    p = "results/H_forcing_convergence/forcing_conv.dat"
    data["forcing_transverse"] = {"file": p, "header": header(p), "rows": rows(p)}

    sl = {}
    for name in ("d2q9_cm", "d3q27_cm"):
        p = f"results/B_shear_layer/shear_{name}.dat"
        sl[name] = {"file": p, "header": header(p), "rows": rows(p)}
    data["shear_layer"] = sl

    cav = {}
    for re_ in (100, 400, 1000):
        for lat in ("d2q9", "d3q27"):
            for tail in ("", "_full"):
                p = f"results/C_cavity/cavity_re{re_}_{lat}_cm{tail}.dat"
                if os.path.exists(os.path.join(REF, p)):
                    cav[f"re{re_}_{lat}{tail}"] = {"file": p, "header": header(p), "rows": rows(p)}
    data["cavity"] = cav

    pin = {}
    for lat in ("d2q9", "d3q27"):
        p = f"results/A_poiseuille_inlet/poiseuille_{lat}_cm.dat"
        pin[lat] = {"file": p, "header": header(p), "rows": rows(p)}
    data["poiseuille_inlet"] = pin

    dip = {}
    for re_ in (625, 2500):
        for lat in ("d2q9", "d3q27"):
            p = f"results/D_dipole/dipole_re{re_}_{lat}_cm.dat"
            dip[f"re{re_}_{lat}"] = {"file": p, "header": header(p), "rows": rows(p)}
    data["dipole"] = dip

    with open(os.path.join(REF, "README.md")) as f:
        lines = f.read().splitlines()

    ln, tab = md_table_after(lines, r"^\| lattice \| operator \| U0=0 \| U0=0.05")
    gal = {}
    for r in tab:
        if r[0] != "D3Q27":
            continue
        gal[r[1]] = [float(v) for v in r[2:6]]
    data["galilean_nu_eff"] = {"file": f"README.md:{ln}", "N": 32, "tau": 0.8, "U": 0.01,
                               "U0": [0.0, 0.05, 0.10, 0.15], "rows": gal}

    ln, tab = md_table_after(lines, r"^\| tau \| BGK \| TRT\(3/16\) \| MRT \| CentralMoments")
    slip = [[float(c) for c in r] for r in tab]
    data["poiseuille_slip"] = {"file": f"README.md:{ln}", "H": 32,
                               "columns": ["tau", "BGK", "TRT", "MRT", "CM"], "rows": slip}

    ln, tab = md_table_after(lines, r"^\| tau \| H \| L2 rel err \| wall slip")
    data["poiseuille_bgk_l2"] = {"file": f"README.md:{ln}",
                                 "columns": ["tau", "H", "L2", "slip"],
                                 "rows": [[float(c) for c in r] for r in tab]}

    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
