"""Regularised velocity walls on the HIP path (FluidSolver.hpp:372-460,
663-710, 793-1033, 1076-1185; boundary/Regularized.hpp), through the C-ABI:
bit-for-bit-class parity with the oracle on cavities, a closed 3-D box and
forced channels (velocity relative L2 <= 1e-12 FP64 / 1e-5 FP32), and the
reference's own steady-state cavity centrelines
(results/C_cavity/cavity_re*_cm*.dat, validation/cavity_cm.cpp, N = 129)."""
import sys

import numpy as np
import pytest

from golden_cases import GOLD, cavity_run, lerp_at
from m3lb_amd import abi, cases
from m3lb_amd.solver import FluidSolver
from oracle.binding import OracleSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def announce(request):
    print(f"\n[gpu-case] {request.node.name}", flush=True)
    sys.stdout.flush()
    yield


def _field(cfg, np_):
    x, y, z, _ = cases.padded_coords(cfg)
    return [1.0 + 0.01 * np.sin(0.5 * x + 0.3 * z), 0.02 * np.cos(0.4 * y),
            0.01 * np.sin(0.3 * x), 0.005 * np.cos(0.2 * x + 0.1 * y) * (cfg.nz > 1)]


def _compare(s, o, dtype, chunks=(1, 1, 38)):
    for chunk in chunks:
        o.step(chunk)
        s.step(chunk)
        ro, uxo, uyo, uzo = o.compute_macroscopic()
        rs, uxs, uys, uzs = s.compute_macroscopic()
        num = np.sum((uxs - uxo) ** 2 + (uys - uyo) ** 2 + (uzs - uzo) ** 2)
        den = np.sum(uxo ** 2 + uyo ** 2 + uzo ** 2)
        l2 = np.sqrt(num / den)
        tol = 1e-12 if dtype == abi.F64 else 1e-5
        assert l2 <= tol, (chunk, l2)
        assert np.abs(rs - ro).max() <= (1e-13 if dtype == abi.F64 else 2e-6)
    gp = s.gather_populations()
    go = o.gather_populations()
    scale = np.abs(go).max()
    assert np.abs(gp - go).max() <= scale * (1e-13 if dtype == abi.F64 else 2e-5)


def _pair(cfg, walls, fd=True, flags=None):
    o = OracleSolver(cfg)
    s = FluidSolver(cfg)
    if flags is not None:
        o.set_geometry(flags)
        s.set_geometry(flags)
    o.set_regularized_walls(*walls)
    s.set_regularized_walls(*walls)
    o.set_fd_corners(fd)
    s.set_fd_corners(fd)
    field = _field(cfg, o.np)
    o.initialize_field(*field)
    s.initialize_field(*field)
    return s, o


OPS = [abi.OP_BGK, abi.OP_BGK_PRODUCT, abi.OP_TRT, abi.OP_TRT_PRODUCT, abi.OP_MRT, abi.OP_CM]


@pytest.mark.parametrize("storage", [abi.SHIFTED, abi.RAW])
@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("lat,nz", [(abi.D2Q9, 1), (abi.D3Q27, 1), (abi.D3Q27, 4)])
def test_cavity_matches_oracle(lat, nz, op, dtype, storage):
    """Lid-driven cavity with corners (finite-difference corner stress)."""
    N = 17
    w = 1.3
    cfg = abi.make_config(lat, op, storage, dtype, (N, N, nz), periodic=(False, False, True),
                          omega=w, omega_minus=1.0 / ((3.0 / 16.0) / (1.0 / w - 0.5) + 0.5),
                          omega_bulk=1.1 if op in (abi.OP_MRT, abi.OP_CM) else -1.0)
    s, o = _pair(cfg, cases.cavity_walls(N, 0.05, nz))
    with s:
        _compare(s, o, dtype)


@pytest.mark.parametrize("fd", [True, False])
@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("op", [abi.OP_BGK, abi.OP_CM])
def test_closed_box_matches_oracle(op, dtype, fd):
    """All six faces walls, edges and vertices corners (density extrapolated
    along a face, stress from one-sided differences on three axes)."""
    shape = (10, 9, 8)
    cfg = abi.make_config(abi.D3Q27, op, abi.SHIFTED, dtype, shape, periodic=(0, 0, 0),
                          omega=1.4)
    s, o = _pair(cfg, cases.box_walls(shape, 0.04), fd=fd)
    with s:
        _compare(s, o, dtype)


@pytest.mark.parametrize("force", ["guo", "field"])
@pytest.mark.parametrize("op", [abi.OP_BGK, abi.OP_TRT, abi.OP_CM])
@pytest.mark.parametrize("lat", [abi.D2Q9, abi.D3Q27])
def test_forced_channel_matches_oracle(lat, op, force):
    """Couette-Poiseuille channel: the Guo odd part in the unknown directions
    and in the rebuilt set, uniform or per node."""
    shape = (12, 11, 1 if lat == abi.D2Q9 else 5)
    cfg = abi.make_config(lat, op, abi.SHIFTED, abi.F64, shape, periodic=(1, 0, 1), omega=0.9,
                          omega_minus=1.1,
                          force=(2e-5, 1e-6, 0.0) if force == "guo" else "field")
    s, o = _pair(cfg, cases.channel_walls(shape, 0.03))
    with s:
        if force == "field":
            x, y, z, _ = cases.padded_coords(cfg)
            ff = (2e-5 * np.sin(0.5 * y + 0.2 * x), 1e-6 * np.cos(0.3 * x), 3e-7 * np.sin(0.4 * z))
            s.set_force_field(*ff)
            o.set_force_field(*ff)
        _compare(s, o, abi.F64)


def test_walls_next_to_a_solid_block():
    """A Solid block touching the floor: directions streaming from it are
    unknown at the wall nodes beside it."""
    N = 16
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64, (N, N, 1),
                          periodic=(False, False, True), omega=1.2)
    fl = np.zeros((1, N, N), dtype=np.uint8)
    fl[0, 1:4, 6:9] = abi.SOLID
    nrm, tag, table = cases.cavity_walls(N, 0.05)
    s, o = _pair(cfg, (nrm, tag, table), flags=fl)
    with s:
        _compare(s, o, abi.F64)


def test_invalid_walls_are_rejected_and_change_nothing():
    N = 12
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64, (N, N, 1),
                          periodic=(False, False, True), omega=1.2)
    nrm, tag, table = cases.cavity_walls(N, 0.05)
    o = OracleSolver(cfg)
    with FluidSolver(cfg) as s:
        bad = nrm.copy()
        bad[5] = 11
        with pytest.raises(RuntimeError):
            s.set_regularized_walls(bad, tag, table)
        badtag = tag.copy()
        badtag[nrm.nonzero()[0][0]] = 7
        with pytest.raises(RuntimeError):
            s.set_regularized_walls(nrm, badtag, table)
        # nothing was applied: the solver runs as the plain (wall-less) box
        o.initialize(1.0, 0.01, 0.0, 0.0)
        s.initialize(1.0, 0.01, 0.0, 0.0)
        o.step(10)
        s.step(10)
        np.testing.assert_allclose(s.gather_populations(), o.gather_populations(), rtol=0,
                                   atol=1e-15)
        # set_geometry removes walls set earlier
        s.set_regularized_walls(nrm, tag, table)
        s.set_geometry(np.zeros((1, N, N), dtype=np.uint8))
        o2 = OracleSolver(cfg)
        o2.initialize(1.0, 0.01, 0.0, 0.0)
        s.initialize(1.0, 0.01, 0.0, 0.0)
        o2.step(10)
        s.step(10)
        np.testing.assert_allclose(s.gather_populations(), o2.gather_populations(), rtol=0,
                                   atol=1e-15)


def _log(msg):
    print(msg, flush=True)


# (file, lattice, Re, restol): the files with full centrelines were written
# with the driver's whole-field rule, "-restol 1e-5 -probe 500"
# (validation/cavity_cm.cpp:81-83, reproduced at its documented 93,500 and
# 321,000 steps) and, for D3Q27 Re = 1000, the tighter 1e-6 of its
# convergence table (:88-91, 528,500 steps); the older 17-station files with
# its default single-node rule (restol 0).
CAVITY = [("re100_d2q9", abi.D2Q9, 100.0, 0.0), ("re400_d2q9", abi.D2Q9, 400.0, 0.0),
          ("re1000_d2q9", abi.D2Q9, 1000.0, 1e-5), ("re100_d3q27", abi.D3Q27, 100.0, 1e-5),
          ("re400_d3q27", abi.D3Q27, 400.0, 1e-5), ("re1000_d3q27", abi.D3Q27, 1000.0, 1e-6)]


@pytest.mark.parametrize("name,lat,Re,restol", CAVITY)
def test_cavity_against_reference_results(name, lat, Re, restol):
    """validation/cavity_cm.cpp at N = 129, u_lid = 0.02, FP64, stepped to the
    stopping rule the file was written with; the 17 Ghia stations (printed
    with 6 decimals) and, where committed, the full centrelines (8 decimals),
    each to its printed rounding."""
    uc, vc, steps = cavity_run(FluidSolver, lat, Re=Re, restol=restol, log=_log)
    print(f"{name}: steady after {steps} steps", flush=True)
    g = GOLD["cavity"][name]
    du = max(abs(lerp_at(uc, r[0]) - r[1]) for r in g["rows"])
    dv = max(abs(lerp_at(vc, r[2]) - r[3]) for r in g["rows"])
    print(f"{name}: 17 stations max |du| {du:.2e} |dv| {dv:.2e}", flush=True)
    assert du <= 5.1e-7 and dv <= 5.1e-7, (du, dv)
    full = GOLD["cavity"].get(name + "_full")
    if full:
        rows = np.array(full["rows"])
        eu = np.abs(uc - rows[:, 1]).max()
        ev = np.abs(vc - rows[:, 2]).max()
        print(f"{name}: full centrelines max |du| {eu:.2e} |dv| {ev:.2e}", flush=True)
        assert eu <= 5.1e-9 and ev <= 5.1e-9, (eu, ev)


# ---------------------------------------------------------------------------
# Outflow codes, specular cells, wall-state controls: device against oracle
# ---------------------------------------------------------------------------
from golden_cases import (plane_fit, poiseuille_inlet_run, specular_channel,  # noqa: E402
                          specular_node_2d, specular_node_3d, specular_node_octant)


def _outflow_walls(shape, code, faces=("x",)):
    """Walls at y = 0 and ny-1 (at rest), a parabolic inlet at x = 0 and the
    outflow code on the x = nx-1 column (and, for faces containing 'ybottom',
    on the floor instead of a wall)."""
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    nrm = np.zeros(x.shape, dtype=np.uint8)
    nrm[x == 0] = cases.NRM_XM
    nrm[x == nx - 1] = code
    nrm[y == ny - 1] = cases.NRM_YP
    nrm[y == 0] = code if "ybottom" in faces else cases.NRM_YM
    tag = np.zeros(x.shape, dtype=np.uint16)
    tag[(x == 0) & (y > 0) & (y < ny - 1)] = 1
    tag[nrm == code] = 2
    table = np.array([[0.0, 0.0, 0.0, 1.0], [0.03, 0.0, 0.0, 1.0], [0.0, 0.0, 0.0, 1.01]])
    return nrm.reshape(-1), tag.reshape(-1), table


OUTFLOW = [(cases.NRM_OUT_XP, 1), (cases.NRM_OUT_XP, 2), (cases.NRM_OUT_EQ, 1),
           (cases.NRM_OUT_FREE, 1)]


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("op", [abi.OP_BGK, abi.OP_CM])
@pytest.mark.parametrize("lat,nz", [(abi.D2Q9, 1), (abi.D3Q27, 3)])
@pytest.mark.parametrize("code,order", OUTFLOW)
def test_outflow_codes_match_oracle(code, order, lat, nz, op, dtype):
    """The outlets read their upstream node after its update in the same
    step (the reference's node order), on the device through the wall pass
    that runs after the step."""
    shape = (14, 9, nz)
    cfg = abi.make_config(lat, op, abi.SHIFTED, dtype, shape, periodic=(False, False, True),
                          omega=1.3, omega_bulk=1.1 if op == abi.OP_CM else -1.0)
    s, o = _pair(cfg, _outflow_walls(shape, code))
    with s:
        for x in (s, o):
            x.set_outflow_order(order)
            x.set_outflow_density(0.99)
        _compare(s, o, dtype)


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
def test_outflow_donor_later_in_the_order(dtype):
    """Free outlets on the floor take their donor from y = 1, a LATER node:
    it is read before its update (the copy taken before the step)."""
    shape = (12, 10, 1)
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, dtype, shape,
                          periodic=(False, False, True), omega=1.2)
    s, o = _pair(cfg, _outflow_walls(shape, cases.NRM_OUT_FREE, faces=("x", "ybottom")))
    with s:
        _compare(s, o, dtype)


@pytest.mark.parametrize("lat,nz", [(abi.D2Q9, 1), (abi.D3Q27, 4)])
def test_corner_stencil_reads_fluid_on_both_sides(lat, nz):
    """A lid segment ending in a corner node inside the box: its stencil
    reads fluid nodes earlier in the order (below it, already updated) and
    later (beside it, not yet)."""
    N = 14
    cfg = abi.make_config(lat, abi.OP_CM, abi.SHIFTED, abi.F64, (N, N, nz),
                          periodic=(False, False, True), omega=1.25)
    nrm, tag, table = cases.cavity_walls(N, 0.05, nz)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(N), np.arange(N), indexing="ij")
    nrm = nrm.reshape(x.shape).copy()
    tag = tag.reshape(x.shape).copy()
    nrm[(y == N - 1) & (x > 6)] = 0          # the lid stops at x = 6 ...
    nrm[(y == N - 1) & (x == 6)] = cases.NRM_CORNER  # ... in a corner node
    nrm[(y == N - 1) & (x == N - 1)] = cases.NRM_XP
    s, o = _pair(cfg, (nrm.reshape(-1), tag.reshape(-1), table))
    with s:
        _compare(s, o, abi.F64)


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
@pytest.mark.parametrize("lat", [abi.D2Q9, abi.D3Q27])
def test_specular_cells_match_oracle(lat, dtype):
    """Ghost specular rows and on-node specular faces (with an edge) beside
    regularised walls and a Solid block, forced."""
    nx, ny, nz = 12, 11, (1 if lat == abi.D2Q9 else 5)
    cfg = abi.make_config(lat, abi.OP_CM, abi.SHIFTED, dtype, (nx, ny, nz),
                          periodic=(True, False, lat == abi.D2Q9), omega=1.3,
                          force=(1e-5, 0.0, 0.0))
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    fl = np.zeros(x.shape, dtype=np.uint8)
    fl[:, 0, :] = abi.SOLID
    fl[:, 3:5, 4:6] = abi.SOLID
    spec = np.zeros(x.shape, dtype=np.uint8)
    spec[:, ny - 1, :] = cases.NRM_YP
    faces = np.zeros(x.shape, dtype=np.uint8)
    faces[:, ny - 2, :] = cases.SPEC_YP
    if nz > 1:
        faces[0, 1:ny - 1, :] |= cases.SPEC_ZM
        faces[nz - 1, 1:ny - 1, :] |= cases.SPEC_ZP
    faces[:, ny - 2, 0:3] = 0
    nrm = np.zeros(x.shape, dtype=np.uint8)
    nrm[:, ny - 2, 0:3] = cases.NRM_YP
    o = OracleSolver(cfg)
    s = FluidSolver(cfg)
    for q in (s, o):
        q.set_geometry(fl)
        q.set_specular_walls(spec)
        q.set_specular_nodes(faces)
        q.set_regularized_walls(nrm.reshape(-1), np.zeros(nrm.size, dtype=np.uint16),
                                np.array([[0.01, 0.0, 0.0, 1.0]]))
    field = _field(cfg, o.np)
    o.initialize_field(*field)
    s.initialize_field(*field)
    with s:
        _compare(s, o, dtype)


def test_wall_state_controls_match_oracle():
    """Velocity scale, wall density, per-node velocity field and its scale,
    and the force correction switched off, on a forced channel with an
    outlet."""
    shape = (14, 9, 1)
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64, shape,
                          periodic=(False, False, True), omega=1.3, force=(2e-6, 0.0, 0.0))
    s, o = _pair(cfg, _outflow_walls(shape, cases.NRM_OUT_XP))
    rng = np.random.default_rng(2)
    u = 0.01 * rng.standard_normal(shape[0] * shape[1])
    with s:
        for q in (s, o):
            q.set_wall_velocity_scale(0.7)
            q.set_wall_density(1.005)
            q.set_force_correction(False)
        _compare(s, o, abi.F64, chunks=(1, 9))
        for q in (s, o):
            q.set_wall_velocity_field(u, 0.3 * u, 0 * u)
            q.set_wall_velocity_field_scale(0.5)
            q.set_force_correction(True)
        _compare(s, o, abi.F64, chunks=(1, 9))


def test_stored_macroscopic_at_boundary_nodes():
    """step(store_macroscopic) writes the state the step collided: at a
    velocity wall that is the imposed velocity and the closure density, at an
    on-node specular node the mirrored set's moments -- what
    compute_macroscopic reports just before the step."""
    nx, ny = 12, 9
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64, (nx, ny, 1),
                          periodic=(True, False, True), omega=1.3)
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    nrm = np.where(y == 0, cases.NRM_YM, 0).astype(np.uint8)
    faces = np.where(y == ny - 1, cases.SPEC_YP, 0).astype(np.uint8)
    with FluidSolver(cfg) as s:
        s.set_regularized_walls(nrm.reshape(-1), np.zeros(nrm.size, dtype=np.uint16),
                                np.array([[0.02, 0.0, 0.0, 1.0]]))
        s.set_specular_nodes(faces)
        s.initialize_field(*_field(cfg, s.np))
        s.step(7)
        before = s.compute_macroscopic()
        s.step(1, store_macroscopic=True)
        stored = s.macroscopic_stored()
        for a, b in zip(before, stored):
            np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-15)


def test_boundary_argument_checks():
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64, (8, 8, 1),
                          periodic=(False, False, True))
    with FluidSolver(cfg) as s:
        with pytest.raises(RuntimeError):
            s.set_wall_velocity_scale(2.0)      # no walls yet
        bad = np.zeros(64, dtype=np.uint8)
        bad[9] = cases.NRM_CORNER
        with pytest.raises(RuntimeError):
            s.set_specular_walls(bad)
        both = np.zeros(64, dtype=np.uint8)
        both[9] = cases.SPEC_YP | cases.SPEC_YM
        with pytest.raises(RuntimeError):
            s.set_specular_nodes(both)
        out = np.zeros(64, dtype=np.uint8)
        out[8 * 3 + 1] = cases.NRM_OUT_XP       # x = 1: no second upstream node
        with pytest.raises(RuntimeError):
            s.set_regularized_walls(out, np.zeros(64, dtype=np.uint16),
                                    np.array([[0.0, 0.0, 0.0, 1.0]]))


@pytest.mark.parametrize("lat,Ly", [(abi.D2Q9, 41), (abi.D2Q9, 81), (abi.D2Q9, 161),
                                    (abi.D3Q27, 41), (abi.D3Q27, 81)])
def test_inlet_poiseuille_against_reference_results(lat, Ly):
    """results/A_poiseuille_inlet/poiseuille_<lat>_cm.dat: error to its 9
    printed digits and the step count of every row run here."""
    name = "d2q9" if lat == abi.D2Q9 else "d3q27"
    row = next(r for r in GOLD["poiseuille_inlet"][name]["rows"] if int(r[0]) == Ly)
    err, steps, _ = poiseuille_inlet_run(FluidSolver, lat, Ly)
    print(f"inlet Poiseuille {name} Ly={Ly}: err {err:.8e} (file {row[5]:.8e}), "
          f"steps {steps} (file {int(row[2])})", flush=True)
    assert steps == int(row[2])
    assert abs(err - row[5]) <= 6e-9 * row[5]


def test_inlet_poiseuille_outflow_order_two_on_device():
    err, _, _ = poiseuille_inlet_run(FluidSolver, abi.D2Q9, 41, oo=2)
    assert abs(err - 1.388e-3) <= 5e-7, err


@pytest.mark.parametrize("lat,tau", [(abi.D2Q9, 0.8), (abi.D2Q9, 1.2), (abi.D3Q27, 0.8)])
def test_specular_ghost_identity_on_device(lat, tau):
    """validation/specular.cpp at its own settings (H = 16, 40000 steps)."""
    H = 16
    uf, _ = specular_channel(FluidSolver, lat, H, tau, False, 40000)
    uh, drift = specular_channel(FluidSolver, lat, H, tau, True, 40000)
    rel = np.abs(uh[1:H + 1] - uf[1:H + 1]).max() / np.abs(uf[1:H + 1]).max()
    assert rel < 1e-12, rel
    lo, vertex = plane_fit(uh, H)
    assert abs(vertex - (H + 0.5)) < 0.02 and abs(lo - 0.5) < 0.02, (lo, vertex)
    assert abs(drift) < 1e-13


@pytest.mark.parametrize("lat", [abi.D2Q9, abi.D3Q27])
def test_specular_node_identity_on_device(lat):
    a, ra = specular_node_2d(FluidSolver, lat, 0)
    b, rb = specular_node_2d(FluidSolver, lat, 1)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    assert np.abs(ra - rb).max() <= 1e-12
    if lat == abi.D3Q27:
        for run in (specular_node_3d, specular_node_octant):
            a, ra = run(FluidSolver, False)
            b, rb = run(FluidSolver, True)
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
            assert np.abs(ra - rb).max() <= 1e-12


from golden_cases import dipole_run  # noqa: E402


@pytest.mark.parametrize("lat", [abi.D2Q9, abi.D3Q27])
@pytest.mark.parametrize("Re", [625, 2500])
def test_dipole_wall_collision_against_reference_results(lat, Re):
    """results/D_dipole/dipole_re<Re>_<lat>_cm.dat (validation/dipole.cpp,
    D_lb = 512, 1025^2 nodes, 19200 steps): E/E0 and Omega/Omega0 at all 31
    probes, to the printed 9 digits."""
    name = f"re{Re}_" + ("d2q9" if lat == abi.D2Q9 else "d3q27")
    want = GOLD["dipole"][name]["rows"]
    got = dipole_run(FluidSolver, lat, float(Re))
    worst = 0.0
    for (t, e, w), row in zip(got, want):
        assert abs(t - row[0]) < 1e-6
        worst = max(worst, abs(e - row[1]) / row[1], abs(w - row[2]) / row[2])
    print(f"dipole {name}: worst relative deviation {worst:.2e}", flush=True)
    assert worst <= 1e-8, worst


def test_outlet_reading_an_earlier_wall_node():
    """A back-pressure outlet whose upstream node is itself a wall node (an
    equilibrium outlet column just before it): the later wall reads the
    earlier one after its update, so the device runs them in two levels."""
    shape = (14, 9, 1)
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64, shape,
                          periodic=(False, False, True), omega=1.3)
    nrm, tag, table = _outflow_walls(shape, cases.NRM_OUT_XP)
    nrm = nrm.reshape(shape[1], shape[0]).copy()
    tag = tag.reshape(shape[1], shape[0]).copy()
    nrm[1:-1, shape[0] - 2] = cases.NRM_OUT_EQ
    tag[1:-1, shape[0] - 2] = 2
    s, o = _pair(cfg, (nrm.reshape(-1), tag.reshape(-1), table))
    with s:
        _compare(s, o, abi.F64)


@pytest.mark.parametrize("dtype,tol", [(abi.F16, 2e-3), (abi.BF16, 2e-2), (abi.I16, 2e-3),
                                       (abi.E2M13, 2e-3)])
def test_walls_and_mirrors_in_16_bit_storage(dtype, tol):
    """Walls, corners and an outlet with 16-bit storage (decoded and
    re-encoded at their per-direction scales) stay within the storage's
    rounding of the FP32 run."""
    N = 24
    shape = (N, N, 1)

    def run(dt):
        cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, dt, shape,
                              periodic=(False, False, True), omega=1.4)
        nrm, tag, table = cases.cavity_walls(N, 0.05)
        nrm = nrm.reshape(N, N).copy()
        nrm[1:-1, N - 1] = cases.NRM_OUT_XP
        tag = tag.reshape(N, N).copy()
        tag[1:-1, N - 1] = 2
        table = np.vstack([table, [[0.0, 0.0, 0.0, 1.0]]])
        table[:, 3] = 1.0
        with FluidSolver(cfg) as s:
            s.set_regularized_walls(nrm.reshape(-1), tag.reshape(-1), table)
            s.initialize(1.0)
            s.step(300)
            _, ux, uy, _ = s.compute_macroscopic()
        return ux, uy

    ux32, uy32 = run(abi.F32)
    ux16, uy16 = run(dtype)
    l2 = np.sqrt(np.sum((ux16 - ux32) ** 2 + (uy16 - uy32) ** 2) / np.sum(ux32 ** 2 + uy32 ** 2))
    print(f"16-bit walls {dtype}: velocity rel-L2 vs FP32 {l2:.2e}", flush=True)
    assert l2 < tol, l2


@pytest.mark.parametrize("dtype,tol", [(abi.F16, 2e-3), (abi.BF16, 2e-2), (abi.E2M13, 2e-3)])
def test_specular_nodes_in_16_bit_storage(dtype, tol):
    """The on-node mirror identity of validation/specular_node.cpp holds in
    16-bit storage to the storage's rounding."""
    a, _ = specular_node_2d(FluidSolver, abi.D2Q9, 0)
    b, _ = specular_node_2d(FluidSolver, abi.D2Q9, 1, dtype=dtype)
    l2 = np.sqrt(np.sum((a - b) ** 2) / np.sum(a ** 2))
    print(f"16-bit specular {dtype}: rel-L2 vs FP64 full box {l2:.2e}", flush=True)
    assert l2 < tol, l2


@pytest.mark.parametrize("dtype", [abi.F64, abi.F32])
def test_every_boundary_kind_together(dtype):
    """A 3-D duct with all boundary kinds at once: a parabolic velocity
    inlet, a back-pressure outlet, a regularised floor with corners, an
    on-node free-slip (specular) ceiling, a Solid block, periodic z, and a
    per-node Guo force."""
    nx, ny, nz = 18, 12, 6
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, (nx, ny, nz),
                          periodic=(False, False, True), omega=1.35, force="field")
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    fl = np.zeros(x.shape, dtype=np.uint8)
    fl[:, 4:7, 7:10] = abi.SOLID
    nrm = np.zeros(x.shape, dtype=np.uint8)
    tag = np.zeros(x.shape, dtype=np.uint16)
    nrm[y == 0] = cases.NRM_YM
    inlet = (x == 0) & (y > 0) & (y < ny - 1)
    nrm[inlet] = cases.NRM_XM
    tag[inlet] = 1 + (y[inlet] - 1)          # one state per row
    nrm[(x == nx - 1) & (y > 0) & (y < ny - 1)] = cases.NRM_OUT_XP
    tag[(x == nx - 1) & (y > 0) & (y < ny - 1)] = ny
    nrm[((x == 0) | (x == nx - 1)) & ((y == 0) | (y == ny - 1))] = cases.NRM_CORNER
    table = np.zeros((ny + 1, 4))
    table[:, 3] = 1.0
    for r in range(1, ny - 1):
        s_ = r / (ny - 1.0)
        table[r, 0] = 0.12 * s_ * (1.0 - s_)
    table[ny, 3] = 0.998
    faces = np.zeros(x.shape, dtype=np.uint8)
    faces[(y == ny - 1) & (x > 0) & (x < nx - 1)] = cases.SPEC_YP
    px, py, pz, _ = cases.padded_coords(cfg)
    ff = (1e-6 * np.sin(0.3 * py), 2e-7 * np.cos(0.2 * px), 1e-7 * np.sin(0.5 * pz))
    o = OracleSolver(cfg)
    s = FluidSolver(cfg)
    for q in (s, o):
        q.set_geometry(fl)
        q.set_specular_nodes(faces)
        q.set_regularized_walls(nrm.reshape(-1), tag.reshape(-1), table)
        q.set_force_field(*ff)
    field = _field(cfg, o.np)
    o.initialize_field(*field)
    s.initialize_field(*field)
    with s:
        _compare(s, o, dtype)
