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


def tgv3d_run(solver_cls, probes, dtype=abi.F64, **kw):
    """validation/tgv3d.cpp:55-104 at its committed configuration: D = 64,
    Re = 1600, u0 = 0.02, CM shifted.  Returns [(t*, E/E0, Psi/Psi0)] for the
    first `probes` probes (every T/20 = 1600 steps)."""
    D, u0, Re, tmax = 64, 0.02, 1600.0, 10.0
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


def momentum_exchange_np(cfg, flags, body, gathered):
    """validation/square_cylinder.cpp:340-350 over every Fluid node: for each
    link (n, i) whose neighbour n + c_i (periodic wrap, else clamped into the
    padded box) is in the body, F += 2 c_i f(n, opp(i)), f = the gathered
    (stored) value unshifted in the compute type (unshift<C>).  flags and body
    are interior arrays, x fastest; gathered is (n_padded, Q)."""
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
    return F
