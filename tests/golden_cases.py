"""Drivers of the reference's validation cases, shared by the oracle tests and
the GPU tests so both run the identical case definition.  `solver_cls` is
oracle.binding.OracleSolver or m3lb_amd.solver.FluidSolver; both expose the
reference's FluidSolver method names."""
import json
import math
import os

import numpy as np

from m3lb_amd import abi, cases

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_data.json")))
OPS = {"BGK": abi.OP_BGK, "TRT": abi.OP_TRT, "MRT": abi.OP_MRT, "CM": abi.OP_CM}
GAL = {"BGK": "BGK", "TRT": "TRT(3/16)", "MRT": "MRT (raw)", "CM": "CentralMoments"}
TG2 = {"d2q9_cm": (abi.D2Q9, abi.OP_CM), "d3q27_bgk": (abi.D3Q27, abi.OP_BGK_PRODUCT),
       "d3q27_cm": (abi.D3Q27, abi.OP_CM), "d3q27_mrt": (abi.D3Q27, abi.OP_MRT)}
# 9 printed significant digits: agreement to 1e-8 relative is the printing limit.
PRINT_TOL = 2e-8


def rel(a, b):
    return abs(a - b) / abs(b)


def magic_minus(w):
    return 1.0 / ((3.0 / 16.0) / (1.0 / w - 0.5) + 0.5)


def decaying(case, lat, op, N, tau=0.8, frac=0.5, U=None, U0=0.0, dtype=abi.F64,
             solver_cls=None, **kw):
    """validation/DecayingFlows.hpp:128-198 + run_and_measure (:93-119).
    Returns (l2_u, nu_eff, steps)."""
    nu = (tau - 0.5) * (1.0 / 3.0)
    k = 2.0 * math.pi / N
    coef = 2.0 * k * k if case == "taylor_green" else k * k
    steps = cases.decaying_steps(nu, coef, frac)
    w = 1.0 / tau
    shape = (N, N, (8 if lat == abi.D3Q27 else 1)) if case == "taylor_green" else (N, N, N)
    cfg = abi.make_config(lat, op, abi.SHIFTED, dtype, shape, omega=w,
                          omega_minus=magic_minus(w))
    s = solver_cls(cfg, **kw)
    if case == "taylor_green":
        U = 0.64 / N if U is None else U
        s.initialize_field(*cases.taylor_green(cfg, U))
    else:
        U = 0.32 / N if U is None else U
        s.initialize_field(*cases.abc_flow(cfg, U, U0))
    idx = cases.interior_index(cfg)

    def norm():
        _, ux, uy, uz = s.compute_macroscopic()
        return math.sqrt(float(np.sum((ux[idx] - U0) ** 2 + (uy[idx] - U0) ** 2 +
                                      (uz[idx] - U0) ** 2)))

    n0 = norm()
    s.step(steps)
    n1 = norm()
    nu_eff = -math.log(n1 / n0) / (coef * steps)
    _, ux, uy, uz = s.compute_macroscopic()
    E = math.exp(-nu * coef * steps)
    ana = (cases.tg_analytic(cfg, U, k, E) if case == "taylor_green"
           else cases.abc_analytic(cfg, U, k, E))
    if U0:
        ana = tuple(a + U0 for a in ana)
    return cases.relative_l2(cfg, ux, uy, uz, ana), nu_eff, steps


def tgv2d_run(lat, op, N, Re=1000.0, u0=0.02, dtype=abi.F64, solver_cls=None, **kw):
    """validation/tgv2d.cpp:69-157: error at t = T against exp(-1) decay."""
    xi = 2.0 * math.pi / N
    nu = u0 * N / Re
    T = int(1.0 / (2.0 * xi * xi * nu))
    cfg = abi.make_config(lat, op, abi.SHIFTED, dtype, (N, N, 1), omega=1.0 / (nu * 3.0 + 0.5))
    s = solver_cls(cfg, **kw)
    s.initialize_field(*cases.tgv2d(cfg, u0))
    s.step(T)
    _, ux, uy, _ = s.compute_macroscopic()
    x, y, _ = cases.interior_xyz(cfg)
    dec = math.exp(-1.0)
    ax = -u0 * np.cos(xi * x) * np.sin(xi * y) * dec
    ay = u0 * np.sin(xi * x) * np.cos(xi * y) * dec
    idx = cases.interior_index(cfg)
    num = float(np.sum((ux[idx] - ax) ** 2 + (uy[idx] - ay) ** 2))
    return math.sqrt(num / float(np.sum(ax * ax + ay * ay))), T


def poiseuille(op, tau, H=32, dtype=abi.F64, storage=abi.SHIFTED, solver_cls=None,
               **kw):
    cfg, fl, nu, G = cases.poiseuille_config(op, tau, H, dtype=dtype, storage=storage)
    s = solver_cls(cfg, **kw)
    s.set_geometry(fl)
    s.initialize(1.0)
    return cases.poiseuille_run(s, cfg, H, nu, G)


def tgv3d_run(solver_cls, probes, dtype=abi.F64, Re=1600.0, **kw):
    """validation/tgv3d.cpp:55-104 at its committed configuration: D = 64,
    Re = 1600 (or the committed 30000), u0 = 0.02, CM shifted.  Returns
    [(t*, E/E0, Psi/Psi0)] for the first `probes` probes (every T/20 = 1600
    steps)."""
    D, u0, tmax = 64, 0.02, 10.0
    nu = u0 * D / Re
    T = int(tmax * D / u0)
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, (D, D, D),
                          omega=1.0 / (nu * 3.0 + 0.5))
    s = solver_cls(cfg, **kw)
    s.initialize_field(*cases.tgv3d_campaign(cfg, u0))
    e0, w0, _ = s.energy_enstrophy()
    out = [(0.0, 1.0, 1.0)]
    probe = T // 20
    for k in range(1, probes + 1):
        s.step(probe)
        e, w, ok = s.energy_enstrophy()
        assert ok
        out.append((tmax * k * probe / T, e / e0, w / w0))
    return out


def momentum_exchange_np(cfg, flags, body, gathered, magnitude=False):
    """validation/square_cylinder.cpp:340-350 over every Fluid node: for each
    link (n, i) whose neighbour n + c_i (periodic wrap, else clamped into the
    padded box) is in the body, F += 2 c_i f(n, opp(i)), f = the gathered
    (stored) value unshifted in the compute type (unshift<C>).  flags and body
    are interior arrays, x fastest; gathered is (n_padded, Q).  magnitude=True
    also returns the sum of |2 c_i f| per axis, the scale of the rounding
    error of F when its link terms cancel."""
    from m3lb_amd import lattice
    d = abi.domain_of(cfg)
    nx, ny, nz, hx, hy, hz, sx, sy, sz, npad = d
    Q = cfg.lattice
    c = np.asarray(lattice.velocities(Q), dtype=np.int64)
    w = np.asarray(lattice.weights(Q))
    ct = np.float64 if cfg.dtype == abi.F64 else np.float32
    shifted = cfg.storage == abi.SHIFTED and cfg.op != abi.OP_NULL
    pf = np.full((sz, sy, sx), abi.EXCLUDED, dtype=np.uint8)
    pb = np.zeros((sz, sy, sx), dtype=bool)
    pf[hz:hz + nz, hy:hy + ny, hx:hx + nx] = np.asarray(flags).reshape(nz, ny, nx)
    pb[hz:hz + nz, hy:hy + ny, hx:hx + nx] = np.asarray(body).reshape(nz, ny, nx) != 0
    zz, yy, xx = np.nonzero(pf == abi.FLUID)
    n = (zz * sy + yy) * sx + xx

    def shift(v, h, m, s, per):
        if per:
            return np.where(v < h, v + m, np.where(v >= h + m, v - m, v))
        return np.clip(v, 0, s - 1)

    F = np.zeros(3)
    M = np.zeros(3)
    per = cfg.periodic
    for i in range(1, Q):
        mx = shift(xx + c[i, 0], hx, nx, sx, per[0])
        my = shift(yy + c[i, 1], hy, ny, sy, per[1])
        mz = shift(zz + c[i, 2] if c.shape[1] > 2 else zz, hz, nz, sz, per[2])
        hit = pb[mz, my, mx]
        if not hit.any():
            continue
        io = lattice.opp(i)
        f = gathered[n[hit], io].astype(ct)
        if shifted:
            f = (f + ct(w[io])).astype(ct)
        f = f.astype(np.float64)
        for a in range(min(3, c.shape[1])):
            F[a] += np.sum(2.0 * c[i, a] * f)
            M[a] += np.sum(np.abs(2.0 * c[i, a] * f))
    return (F, M) if magnitude else F


def forcing_transverse(solver_cls, orient, tau, H, umax=0.02, cap=600000, dtype=abi.F64):
    """validation/forcing_transverse.cpp:53-126: D2Q9 central moments, Guo force
    along the channel, halfway bounce-back walls.  orient 0: walls y-normal,
    force along x; orient 1: walls x-normal, force along y.  Steps in probes of
    200 until the centre velocity changes by less than 1e-15 relative, then
    returns (L2 against the exact parabola, wall slip 0.5 - lower root)."""
    across, along = H + 2, 8
    nx, ny = (across, along) if orient else (along, across)
    nu = (tau - 0.5) / 3.0
    G = 8.0 * 1.0 * nu * umax / (H * H)
    force = (0.0, G, 0.0) if orient else (G, 0.0, 0.0)
    cfg = abi.make_config(abi.D2Q9, abi.OP_CM, abi.SHIFTED, dtype, (nx, ny, 1),
                          periodic=(orient == 0, orient == 1, True), omega=1.0 / tau,
                          force=force)
    fl = np.zeros((ny, nx), dtype=np.uint8)
    if orient:
        fl[:, 0] = fl[:, across - 1] = abi.SOLID
    else:
        fl[0, :] = fl[across - 1, :] = abi.SOLID
    d = abi.domain_of(cfg)
    sx, hx, hy = d[6], d[3], d[4]
    pid = lambda x, y: (y + hy) * sx + x + hx
    s = solver_cls(cfg)
    try:
        s.set_geometry(fl.reshape(-1))
        s.initialize(1.0)
        comp = 2 if orient else 1
        prev = 0.0
        for t in range(0, cap, 200):
            s.step(200)
            cur = s.compute_macroscopic()[comp][pid(nx // 2, ny // 2)]
            if t > 0 and abs(cur - prev) < 1e-15 * abs(cur):
                break
            prev = cur
        m = s.compute_macroscopic()
    finally:
        s.close()
    v, rho = m[comp], m[0]
    ids = [pid(j, ny // 2) if orient else pid(nx // 2, j) for j in range(1, H + 1)]
    rho_mean = sum(rho[i] for i in ids) / H
    A = G / (2.0 * rho_mean * nu)
    num = den = 0.0
    for j, i in zip(range(1, H + 1), ids):
        ana = A * (j - 0.5) * (H + 0.5 - j)
        num += (v[i] - ana) ** 2
        den += ana * ana
    lo, _ = cases.parabola_roots([float(j) for j in range(1, H + 1)], [v[i] for i in ids])
    return math.sqrt(num / den), 0.5 - lo


def forcing_3d_channel(solver_cls, wall, fdir, tau, H=16, umax=0.02, cap=4000000,
                       dtype=abi.F64):
    """validation/forcing_3d.cpp:93-150 (PART 1): D3Q27 central moments, walls
    normal to axis `wall`, Guo force along axis `fdir`, periodic elsewhere,
    8 nodes along the other axes.  Probes every 200 steps until the centre
    velocity changes by less than 1e-15 relative; returns the L2 error against
    u(s) = G / (2 rho nu) (s - 0.5)(H + 0.5 - s)."""
    n = [8, 8, 8]
    n[wall] = H + 2
    nu = (tau - 0.5) / 3.0
    G = 8.0 * nu * umax / (H * H)
    F = [0.0, 0.0, 0.0]
    F[fdir] = G
    per = [True, True, True]
    per[wall] = False
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, tuple(n), periodic=per,
                          omega=1.0 / tau, force=tuple(F))
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    c = (x, y, z)[wall]
    flags = np.where((c == 0) | (c == n[wall] - 1), abi.SOLID, abi.FLUID).astype(np.uint8)
    d = abi.domain_of(cfg)
    hx, hy, hz, sx, sy = d[3], d[4], d[5], d[6], d[7]
    pid = lambda p: ((p[2] + hz) * sy + (p[1] + hy)) * sx + p[0] + hx

    def at(j):
        p = [n[0] // 2, n[1] // 2, n[2] // 2]
        p[wall] = j
        return pid(p)

    s = solver_cls(cfg)
    try:
        s.set_geometry(flags.reshape(-1))
        s.initialize(1.0)
        comp = 1 + fdir
        prev = 0.0
        for t in range(0, cap, 200):
            s.step(200)
            cur = s.compute_macroscopic()[comp][at(n[wall] // 2)]
            if t > 0 and abs(cur - prev) < 1e-15 * abs(cur):
                break
            prev = cur
        m = s.compute_macroscopic()
    finally:
        s.close()
    ids = [at(j) for j in range(1, H + 1)]
    rho_mean = sum(m[0][i] for i in ids) / H
    A = G / (2.0 * rho_mean * nu)
    num = den = 0.0
    for j, i in zip(range(1, H + 1), ids):
        ana = A * (j - 0.5) * (H + 0.5 - j)
        num += (m[comp][i] - ana) ** 2
        den += ana * ana
    return math.sqrt(num / den)


def forcing_3d_sine(solver_cls, N, tau, umax=0.02, cap=4000000, dtype=abi.F64, threads=None):
    """validation/forcing_3d.cpp:152-200 (PART 2): periodic cube, D3Q27 central
    moments, per-node force Fx = F0 sin(k y) sin(k z) (SinForce, :64-91, with
    y, z the interior coordinates), whose steady state ux = umax sin(k y)
    sin(k z) is an exact Navier-Stokes solution.  Probes every 100 steps until
    ux at (0, N/4, N/4) changes by less than 1e-14 relative; returns (L2 of the
    velocity error, peak |ux| / umax - 1)."""
    nu = (tau - 0.5) / 3.0
    k = 2.0 * math.pi / N
    F0 = umax * 2.0 * nu * k * k
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, dtype, (N, N, N), omega=1.0 / tau,
                          force="field")
    x, y, z, d = cases.padded_coords(cfg)
    fx = F0 * np.sin(k * (y - d[4])) * np.sin(k * (z - d[5]))
    zero = np.zeros_like(fx)
    s = solver_cls(cfg) if threads is None else solver_cls(cfg, threads=threads)
    try:
        s.set_force_field(fx, zero, zero)
        s.initialize(1.0)
        probe = (N // 4 * N + N // 4) * N
        prev = 0.0
        for t in range(0, cap, 100):
            s.step(100)
            cur = s.compute_macroscopic()[1][probe]
            if t > 0 and abs(cur - prev) < 1e-14 * (abs(cur) + 1e-30):
                break
            prev = cur
        _, ux, uy, uz = s.compute_macroscopic()
    finally:
        s.close()
    ana = umax * np.sin(k * (y - d[4])) * np.sin(k * (z - d[5]))
    num = float(np.sum((ux - ana) ** 2 + uy ** 2 + uz ** 2))
    den = float(np.sum(ana ** 2))
    return math.sqrt(num / den), float(np.max(np.abs(ux))) / umax - 1.0


def shear_layer_run(solver_cls, lattice, probes, dtype=abi.F64, **kw):
    """validation/shear_layer.cpp:23-80 at its committed configuration: L = 256,
    Re = 3e4, u0 = 0.02, CM shifted (Campaign.hpp OpCM), probes every
    t0/40 = 320 steps.  Returns [(t/t0, E/E0, Psi/Psi0)] for probes 0..probes."""
    L, Re, u0 = 256, 3e4, 0.02
    nu = u0 * L / Re
    t0 = int(L / u0)
    probe = t0 // 40
    cfg = abi.make_config(lattice, abi.OP_CM, abi.SHIFTED, dtype, (L, L, 1),
                          omega=1.0 / (nu * 3.0 + 0.5))
    s = solver_cls(cfg, **kw)
    try:
        s.initialize_field(*cases.shear_layer(cfg, u0))
        e0, w0, _ = s.energy_enstrophy()
        out = [(0.0, 1.0, 1.0)]
        for k in range(1, probes + 1):
            s.step(probe)
            e, w, ok = s.energy_enstrophy()
            assert ok
            out.append((k * probe / t0, e / e0, w / w0))
    finally:
        s.close()
    return out


def cavity_run(solver_cls, lattice, N=129, Re=100.0, U=0.02, dtype=abi.F64, probe=500,
               cap=4_000_000, rel=1e-11, restol=0.0, log=None, **kw):
    """validation/cavity_cm.cpp:180-290: an N x N lid-driven cavity (periodic
    z, nz = 1), regularised walls on the edge nodes (lid at (U, 0, 0), corners
    at rest), CM shifted, nu = U (N - 1) / Re, started from rest at rho = 1,
    stepped in blocks of `probe`.  Stops when the centre node's u_x changes by
    less than rel * (|u| + 1e-12) between two probes (the default), or with
    restol > 0 (its -restol option, :230-246) when the relative L2 change of
    the whole velocity field over one probe falls below restol.  Returns
    (u/U along x = N/2, v/U along y = N/2, steps taken)."""
    nu = U * (N - 1) / Re
    cfg = abi.make_config(lattice, abi.OP_CM, abi.SHIFTED, dtype, (N, N, 1),
                          periodic=(False, False, True), omega=1.0 / (3.0 * nu + 0.5))
    nrm, tag, table = cases.cavity_walls(N, U)
    s = solver_cls(cfg, **kw)
    try:
        s.set_regularized_walls(nrm, tag, table)
        s.initialize(1.0)
        taken, prev = 0, 0.0
        pu = pv = None
        for t in range(0, cap, probe):
            s.step(probe)
            taken += probe
            _, ux, uy, _ = s.compute_macroscopic()
            ux = interior2d(cfg, ux)
            uy = interior2d(cfg, uy)
            cur = ux[N // 2, N // 2]
            assert math.isfinite(cur), f"diverged at step {taken}"
            if log and taken % 50_000 == 0:
                log(f"cavity N={N} Re={Re:g}: step {taken}, centre u {cur:.12e}")
            if restol > 0.0:
                if pu is None:
                    pu, pv = np.zeros_like(ux), np.zeros_like(uy)
                num = float(np.sum((ux - pu) ** 2 + (uy - pv) ** 2))
                den = float(np.sum(ux * ux + uy * uy))
                res = math.sqrt(num / max(den, 1e-300))
                pu, pv = ux, uy
                if t > 0 and res < restol:
                    break
            elif t > 0 and abs(cur - prev) < rel * (abs(cur) + 1e-12):
                break
            prev = cur
        return ux[:, N // 2] / U, uy[N // 2, :] / U, taken
    finally:
        s.close()


def interior2d(cfg, a):
    """(ny, nx) interior of a padded per-node array of a 2-D (nz = 1) box."""
    nx, ny, nz, hx, hy, hz, sx, sy, sz, _ = abi.domain_of(cfg)
    return np.asarray(a).reshape(sz, sy, sx)[hz, hy:hy + ny, hx:hx + nx]


def lerp_at(v, s):
    """validation/cavity_cm.cpp:161-166: linear interpolation of a profile
    sampled at k / (len - 1)."""
    p = s * (len(v) - 1)
    i = int(p)
    if i >= len(v) - 1:
        return float(v[-1])
    return float(v[i] + (p - i) * (v[i + 1] - v[i]))


def specular_channel(solver_cls, lattice, H, tau, spec, steps, **kw):
    """validation/specular.cpp:38-70: a Guo-forced BGK channel (second-order
    equilibrium, shifted), Solid rows at y = 0 and ny - 1; the full channel
    (ny = 2H + 2) or its lower half (ny = H + 2) with the top row a specular
    ghost (NrmYp).  Same force in both.  Returns (u_x(y) at x = nx/2,
    z = nz/2, relative mass drift)."""
    ny = H + 2 if spec else 2 * H + 2
    nx, nz = 6, (6 if lattice == abi.D3Q27 else 1)
    nu = (tau - 0.5) / 3.0
    umax = 0.02
    G = 8.0 * nu * umax / (4.0 * H * H)
    cfg = abi.make_config(lattice, abi.OP_BGK, abi.SHIFTED, abi.F64, (nx, ny, nz),
                          periodic=(True, False, True), omega=1.0 / tau, force=(G, 0.0, 0.0))
    s = solver_cls(cfg, **kw)
    try:
        fl = np.zeros((nz, ny, nx), dtype=np.uint8)
        fl[:, 0, :] = abi.SOLID
        fl[:, ny - 1, :] = abi.SOLID
        s.set_geometry(fl)
        if spec:
            nrm = np.zeros((nz, ny, nx), dtype=np.uint8)
            nrm[:, ny - 1, :] = cases.NRM_YP
            s.set_specular_walls(nrm)
        s.initialize(1.0)
        m0 = s.total_mass()
        s.step(steps)
        drift = (s.total_mass() - m0) / m0
        ux = s.compute_macroscopic()[1]
        nxp, nyp, nzp, hx, hy, hz, sx, sy, sz, _ = abi.domain_of(cfg)
        u = np.asarray(ux).reshape(sz, sy, sx)[hz + nz // 2, hy:hy + ny, hx + nx // 2]
    finally:
        s.close()
    return u, drift


def plane_fit(u, H):
    """validation/specular.cpp:95-121: a quadratic through u(y), y = 1..H;
    returns (lower root, vertex) = (no-slip plane, symmetry plane)."""
    y = np.arange(1, H + 1, dtype=np.float64)
    c2, b, a = np.polyfit(y, np.asarray(u[1:H + 1]), 2)
    disc = b * b - 4 * c2 * a
    r1 = (-b + math.sqrt(disc)) / (2 * c2)
    r2 = (-b - math.sqrt(disc)) / (2 * c2)
    return min(r1, r2), -b / (2 * c2)


def specular_node_2d(solver_cls, lattice, mode, M=16, ny=20, steps=300, tau=0.6, U=0.04,
                     a=0.02, dtype=abi.F64, **kw):
    """validation/specular_node.cpp:38-96: a solenoidal mode odd/even about
    x = 0 and x = M, either on the periodic box of width 2M (mode 0) or on
    [0, M] with on-node specular nodes (mode 1, SpecXm / SpecXp) or specular
    ghosts (mode 2) at x = 0 and M.  BGK second-order, shifted.  Returns
    (u_x, u_y interleaved over x = 0..M, y = 0..ny-1; rho)."""
    half = mode != 0
    nx = M + 1 if half else 2 * M
    cfg = abi.make_config(lattice, abi.OP_BGK, abi.SHIFTED, dtype, (nx, ny, 1),
                          periodic=(not half, True, True), omega=1.0 / tau)
    s = solver_cls(cfg, **kw)
    try:
        if mode == 1:
            f = np.zeros((1, ny, nx), dtype=np.uint8)
            f[:, :, 0] = cases.SPEC_XM
            f[:, :, nx - 1] = cases.SPEC_XP
            s.set_specular_nodes(f)
        elif mode == 2:
            f = np.zeros((1, ny, nx), dtype=np.uint8)
            f[:, :, 0] = cases.NRM_XM
            f[:, :, nx - 1] = cases.NRM_XP
            s.set_specular_walls(f)
        px, py, pz, d = cases.padded_coords(cfg)
        x = (px - d[3]).astype(np.float64)
        y = (py - d[4]).astype(np.float64)
        ux = -U * np.sin(math.pi * x / M) * np.sin(2 * math.pi * y / ny)
        uy = -U * (ny / (2.0 * M)) * np.cos(math.pi * x / M) * np.cos(2 * math.pi * y / ny)
        rr = 1.0 + a * np.cos(math.pi * x / M) * np.cos(2 * math.pi * y / ny)
        s.initialize_field(rr, ux, uy, np.zeros_like(ux))
        s.step(steps)
        r, u, v, _ = (interior2d(cfg, q) for q in s.compute_macroscopic())
    finally:
        s.close()
    uv = np.stack([u[:, :M + 1], v[:, :M + 1]], axis=-1).reshape(-1)
    return uv, r[:, :M + 1].reshape(-1)


def specular_node_3d(solver_cls, quarter, M=16, ny=20, P=12, steps=300, tau=0.6, U=0.04,
                     a=0.02, **kw):
    """validation/specular_node.cpp:98-160, D3Q27: the full periodic box
    (2M x ny x 2P) or its quarter [0, M] x [0, P] with on-node specular faces
    at x = 0, M and z = 0, P (edges carry two faces).  Returns (u, v, w over
    the quarter; rho)."""
    nx = M + 1 if quarter else 2 * M
    nz = P + 1 if quarter else 2 * P
    cfg = abi.make_config(abi.D3Q27, abi.OP_BGK, abi.SHIFTED, abi.F64, (nx, ny, nz),
                          periodic=(not quarter, True, not quarter), omega=1.0 / tau)
    s = solver_cls(cfg, **kw)
    try:
        if quarter:
            z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            f = np.zeros(x.shape, dtype=np.uint8)
            f |= np.where(x == 0, cases.SPEC_XM, 0).astype(np.uint8)
            f |= np.where(x == nx - 1, cases.SPEC_XP, 0).astype(np.uint8)
            f |= np.where(z == 0, cases.SPEC_ZM, 0).astype(np.uint8)
            f |= np.where(z == nz - 1, cases.SPEC_ZP, 0).astype(np.uint8)
            s.set_specular_nodes(f)
        px, py, pz, d = cases.padded_coords(cfg)
        x = (px - d[3]).astype(np.float64)
        y = (py - d[4]).astype(np.float64)
        z = (pz - d[5]).astype(np.float64)
        sc = U / max(M, P)
        A, B, C = sc * M, -sc * ny * 0.5, sc * P
        sx, cx = np.sin(math.pi * x / M), np.cos(math.pi * x / M)
        sz, cz = np.sin(math.pi * z / P), np.cos(math.pi * z / P)
        sy, cy = np.sin(2 * math.pi * y / ny), np.cos(2 * math.pi * y / ny)
        s.initialize_field(1.0 + a * cx * cy * cz, A * sx * cy * cz, B * cx * sy * cz,
                           C * cx * cy * sz)
        s.step(steps)
        mac = s.compute_macroscopic()
        nxp, nyp, nzp, hx, hy, hz, sxp, syp, szp, _ = abi.domain_of(cfg)
        cut = [np.asarray(q).reshape(szp, syp, sxp)[hz:hz + P + 1, hy:hy + ny, hx:hx + M + 1]
               for q in mac]
    finally:
        s.close()
    return np.stack(cut[1:], axis=-1).reshape(-1), cut[0].reshape(-1)


def poiseuille_inlet_run(solver_cls, lattice, Ly, oo=1, U0=0.005, Re=10.0, Lx=21, ctol=1e-10,
                         **kw):
    """validation/poiseuille_inlet.cpp:118-205 with its defaults: a channel of
    Lx x Ly (x, y walls, z periodic, Lz = 1), regularised walls at y = 0 and
    Ly - 1, a parabolic velocity inlet at x = 0 (NrmXm, u = 4 U0 s (1 - s)),
    a constant back-pressure outlet at x = Lx - 1 (NrmOutXp, rho = 1) with
    tangential order `oo`; CM shifted, nu = U0 Ly / Re, from rest; the
    whole-field u_x change over probe = max(500, tau_d / 200) steps below ctol
    after at least tau_d / 2 steps, tau_d = Re Ly / U0.  Returns (relative L2
    error of u_x(x = Lx/2) against the parabola, steps, residual)."""
    nu = U0 * Ly / Re
    cfg = abi.make_config(lattice, abi.OP_CM, abi.SHIFTED, abi.F64, (Lx, Ly, 1),
                          periodic=(False, False, True), omega=1.0 / (3.0 * nu + 0.5))
    z, y, x = np.meshgrid(np.arange(1), np.arange(Ly), np.arange(Lx), indexing="ij")
    nrm = np.zeros(x.shape, dtype=np.uint8)
    vel = np.zeros(x.shape + (4,))
    vel[..., 3] = 1.0  # WallSpec::rho defaults to 1 (FluidSolver.hpp:100-102)
    outlet = (x == Lx - 1) & (y > 0) & (y < Ly - 1)
    nrm[x == 0] = cases.NRM_XM
    nrm[outlet] = cases.NRM_OUT_XP
    nrm[y == 0] = cases.NRM_YM
    nrm[y == Ly - 1] = cases.NRM_YP
    yy = y / (Ly - 1.0)
    vel[..., 0] = np.where(x == 0, 4.0 * U0 * yy * (1.0 - yy), 0.0)
    states = {}
    tag = np.zeros(x.shape, dtype=np.uint16)
    for k in zip(*np.nonzero(nrm)):
        key = tuple(vel[k])
        tag[k] = states.setdefault(key, len(states))
    table = np.array(sorted(states, key=states.get))
    s = solver_cls(cfg, **kw)
    try:
        s.set_regularized_walls(nrm.reshape(-1), tag.reshape(-1), table)
        s.set_outflow_order(oo)
        s.initialize(1.0)
        tau_d = Re * Ly / U0
        probe = max(500, int(tau_d / 200.0))
        tmin = int(0.5 * tau_d)
        cap = int(30.0 * tau_d)
        prev = np.zeros((Ly, Lx))
        taken, res = 0, float("nan")
        for _ in range(0, cap, probe):
            s.step(probe)
            taken += probe
            u = interior2d(cfg, s.compute_macroscopic()[1])
            assert np.all(np.isfinite(u))
            res = math.sqrt(float(np.sum((u - prev) ** 2)) / max(float(np.sum(u * u)), 1e-300))
            prev = u
            if taken >= tmin and res < ctol:
                break
        u = interior2d(cfg, s.compute_macroscopic()[1])[:, Lx // 2]
    finally:
        s.close()
    s_ = np.arange(Ly) / (Ly - 1.0)
    ana = 4.0 * U0 * s_ * (1.0 - s_)
    return math.sqrt(float(np.sum((u - ana) ** 2) / np.sum(ana ** 2))), taken, res


def dipole_run(solver_cls, lattice, Re, Dlb=512, tmax=0.75, u0=0.02, dtype=abi.F64, **kw):
    """validation/dipole.cpp:28-140: the dipole-wall collision in the box
    [-1, 1]^2 of N = 2 Dlb + 1 nodes per side, regularised no-slip walls on
    the edge nodes (corners marked corner), CM shifted, nu = u0 Dlb / Re,
    the two Gaussian monopoles (|we| = 299.56, r0 = 0.1, at y = +-0.1) scaled
    by u0, T = tmax Dlb / u0 steps, E and Omega (Campaign diagnostics) every
    T / 30 steps.  Returns [(t, E/E0, Omega/Omega0)]."""
    N = 2 * Dlb + 1
    nu = u0 * Dlb / Re
    T = int(tmax * Dlb / u0)
    dx = 1.0 / Dlb
    cfg = abi.make_config(lattice, abi.OP_CM, abi.SHIFTED, dtype, (N, N, 1),
                          periodic=(False, False, True), omega=1.0 / (3.0 * nu + 0.5))
    z, y, x = np.meshgrid(np.arange(1), np.arange(N), np.arange(N), indexing="ij")
    xl, xr, yb, yt = x == 0, x == N - 1, y == 0, y == N - 1
    nrm = np.zeros(x.shape, dtype=np.uint8)
    nrm[yt] = cases.NRM_YP
    nrm[yb] = cases.NRM_YM
    nrm[xr] = cases.NRM_XP
    nrm[xl] = cases.NRM_XM
    nrm[(xl | xr) & (yb | yt)] = cases.NRM_CORNER
    s = solver_cls(cfg, **kw)
    try:
        s.set_regularized_walls(nrm.reshape(-1), np.zeros(nrm.size, dtype=np.uint16),
                                np.array([[0.0, 0.0, 0.0, 1.0]]))
        px, py, _, d = cases.padded_coords(cfg)
        X = ((px - d[3]) - Dlb) * dx
        Y = ((py - d[4]) - Dlb) * dx
        we, r0, y1, y2 = 299.56, 0.1, 0.1, -0.1
        ea = np.exp(-(X * X + (Y - y1) ** 2) / (r0 * r0))
        eb = np.exp(-(X * X + (Y - y2) ** 2) / (r0 * r0))
        ux = (-0.5 * we * (Y - y1) * ea + 0.5 * we * (Y - y2) * eb) * u0
        uy = (0.5 * we * X * ea - 0.5 * we * X * eb) * u0
        s.initialize_field(np.ones_like(ux), ux, uy, np.zeros_like(ux))
        e0, w0, _ = s.energy_enstrophy()
        probe = T // 30
        out = [(0.0, 1.0, 1.0)]
        for k in range(1, 31):
            s.step(probe)
            e, w, ok = s.energy_enstrophy()
            assert ok, f"diverged at probe {k}"
            out.append((tmax * k * probe / T, e / e0, w / w0))
    finally:
        s.close()
    return out


def specular_node_octant(solver_cls, octant, M=12, N=10, P=8, steps=300, tau=0.6, U=0.03,
                         a=0.02, dtype=abi.F64, **kw):
    """The three-axis counterpart of validation/specular_node.cpp's quarter
    box: a solenoidal mode with u_x odd in x, u_y odd in y, u_z odd in z (each
    even in the other two) and rho even in all, on the periodic box 2M x 2N
    x 2P or on its octant [0, M] x [0, N] x [0, P] with on-node specular
    faces on all six sides (edges two faces, vertices three).  D3Q27 BGK
    shifted.  Returns (u, v, w over the octant; rho)."""
    nx, ny, nz = (M + 1, N + 1, P + 1) if octant else (2 * M, 2 * N, 2 * P)
    per = not octant
    cfg = abi.make_config(abi.D3Q27, abi.OP_BGK, abi.SHIFTED, dtype, (nx, ny, nz),
                          periodic=(per, per, per), omega=1.0 / tau)
    s = solver_cls(cfg, **kw)
    try:
        if octant:
            z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            f = np.zeros(x.shape, dtype=np.uint8)
            for cond, bit in ((x == 0, cases.SPEC_XM), (x == nx - 1, cases.SPEC_XP),
                              (y == 0, cases.SPEC_YM), (y == ny - 1, cases.SPEC_YP),
                              (z == 0, cases.SPEC_ZM), (z == nz - 1, cases.SPEC_ZP)):
                f |= np.where(cond, bit, 0).astype(np.uint8)
            s.set_specular_nodes(f)
        px, py, pz, d = cases.padded_coords(cfg)
        X = math.pi * (px - d[3]) / M
        Y = math.pi * (py - d[4]) / N
        Z = math.pi * (pz - d[5]) / P
        A, B = U * M, U * N
        C = -(A / M + B / N) * P  # divergence-free
        s.initialize_field(1.0 + a * np.cos(X) * np.cos(Y) * np.cos(Z),
                           A * np.sin(X) * np.cos(Y) * np.cos(Z) / max(M, N, P),
                           B * np.cos(X) * np.sin(Y) * np.cos(Z) / max(M, N, P),
                           C * np.cos(X) * np.cos(Y) * np.sin(Z) / max(M, N, P))
        s.step(steps)
        mac = s.compute_macroscopic()
        nxp, nyp, nzp, hx, hy, hz, sxp, syp, szp, _ = abi.domain_of(cfg)
        cut = [np.asarray(q).reshape(szp, syp, sxp)[hz:hz + P + 1, hy:hy + N + 1, hx:hx + M + 1]
               for q in mac]
    finally:
        s.close()
    return np.stack(cut[1:], axis=-1).reshape(-1), cut[0].reshape(-1)
