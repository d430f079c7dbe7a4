"""Initial conditions and measurements of the reference's validation cases.

Each builder returns padded (n_padded,) arrays rho, ux, uy, uz for
initialize_field, computed in float64 from the same formulas, evaluated at the
same padded coordinates, as the reference's device lambdas.  Measurement
helpers restate the reference's host loops over the macroscopic arrays.
"""
import math

import numpy as np

from . import abi


def padded_coords(cfg):
    """Padded (px, py, pz) of every node, x fastest (src/grid/Domain.hpp:66-70)."""
    d = abi.domain_of(cfg)
    sx, sy, sz = d[6], d[7], d[8]
    n = np.arange(d[9], dtype=np.int64)
    return n % sx, (n // sx) % sy, n // (sx * sy), d


def interior_index(cfg):
    """Padded index of every interior node in x-fastest interior order."""
    d = abi.domain_of(cfg)
    nx, ny, nz, hx, hy, hz, sx, sy = d[:8]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (((z + hz) * sy + (y + hy)) * sx + (x + hx)).reshape(-1)


def taylor_green(cfg, U):
    """Taylor-Green, z-independent (validation/DecayingFlows.hpp:128-159).

    Uses the PADDED coordinates, as the reference lambda does."""
    x, y, _, d = padded_coords(cfg)
    k = 2.0 * math.pi / cfg.nx
    X, Y = k * x, k * y
    rho = 1.0 - 0.25 * U * U * (np.cos(2.0 * X) + np.cos(2.0 * Y)) * 3.0
    ux = -U * np.cos(X) * np.sin(Y)
    uy = U * np.sin(X) * np.cos(Y)
    return rho, ux, uy, np.zeros_like(ux)


def taylor_green_exact(cfg, U, nu, steps):
    k = 2.0 * math.pi / cfg.nx
    E = math.exp(-nu * 2.0 * k * k * steps)
    return k, E


def abc_flow(cfg, U, U0=0.0):
    """ABC/Beltrami flow (validation/DecayingFlows.hpp:164-198), with the uniform
    translation U0 of validation/galilean.cpp:56-69."""
    x, y, z, d = padded_coords(cfg)
    k = 2.0 * math.pi / cfg.nx
    X, Y, Z = k * x, k * y, k * z
    ax = U * (np.sin(Z) + np.cos(Y))
    ay = U * (np.sin(X) + np.cos(Z))
    az = U * (np.sin(Y) + np.cos(X))
    rho = 1.0 - 0.5 * (ax * ax + ay * ay + az * az) * 3.0
    return rho, ax + U0, ay + U0, az + U0


def tgv3d_campaign(cfg, u0):
    """validation/tgv3d.cpp:61-72: rho = 1, three-component Taylor-Green."""
    x, y, z, d = padded_coords(cfg)
    k = 2.0 * math.pi / cfg.nx
    X, Y, Z = k * (x - d[3]), k * (y - d[4]), k * (z - d[5])
    ux = u0 * np.cos(X) * np.sin(Y) * np.sin(Z)
    uy = -0.5 * u0 * np.sin(X) * np.cos(Y) * np.sin(Z)
    uz = -0.5 * u0 * np.sin(X) * np.sin(Y) * np.cos(Z)
    return np.ones_like(ux), ux, uy, uz


def tgv3d_benchmark(cfg, u0, perturbation=0.0, seed=0):
    """validation/tgv3d_bench.cpp:76-90: w = 0 and the pressure-carrying density,
    plus an optional seeded uniform perturbation U(-a, a) per component per
    interior node, drawn in x-fastest interior order (no counterpart in the
    reference; it is part of the benchmark workload definition)."""
    x, y, z, d = padded_coords(cfg)
    k = 2.0 * math.pi / cfg.nx
    X, Y, Z = k * (x - d[3]), k * (y - d[4]), k * (z - d[5])
    dr = 3.0 * u0 * u0 / 16.0
    rho = 1.0 + dr * (np.cos(2.0 * X) + np.cos(2.0 * Y)) * (np.cos(2.0 * Z) + 2.0)
    ux = u0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    uy = -u0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    uz = np.zeros_like(ux)
    if perturbation:
        rng = np.random.Generator(np.random.MT19937(seed))
        idx = interior_index(cfg)
        p = rng.uniform(-perturbation, perturbation, size=(3, idx.size))
        ux[idx] += p[0]
        uy[idx] += p[1]
        uz[idx] += p[2]
    return rho, ux, uy, uz


def tgv2d(cfg, u0):
    """validation/tgv2d.cpp:80-90."""
    x, y, _, d = padded_coords(cfg)
    xi = 2.0 * math.pi / cfg.nx
    X, Y = x - d[3], y - d[4]
    rho = 1.0 - 0.75 * u0 * u0 * (np.cos(2.0 * xi * X) + np.cos(2.0 * xi * Y))
    ux = -u0 * np.cos(xi * X) * np.sin(xi * Y)
    uy = u0 * np.sin(xi * X) * np.cos(xi * Y)
    return rho, ux, uy, np.zeros_like(ux)


def velocity_norm(cfg, ux, uy, uz):
    """validation/DecayingFlows.hpp:72-90."""
    idx = interior_index(cfg)
    a, b, c = ux[idx], uy[idx], uz[idx]
    return math.sqrt(float(np.sum(a * a + b * b + c * c)))


def relative_l2(cfg, ux, uy, uz, ana):
    """Relative L2 of the velocity against an analytic field given on interior
    nodes in x-fastest order (validation/DecayingFlows.hpp:105-118)."""
    idx = interior_index(cfg)
    e0 = ux[idx] - ana[0]
    e1 = uy[idx] - ana[1]
    e2 = uz[idx] - ana[2]
    num = float(np.sum(e0 * e0 + e1 * e1 + e2 * e2))
    den = float(np.sum(ana[0] * ana[0] + ana[1] * ana[1] + ana[2] * ana[2]))
    return math.sqrt(num / den)


def interior_xyz(cfg):
    z, y, x = np.meshgrid(np.arange(cfg.nz), np.arange(cfg.ny), np.arange(cfg.nx),
                          indexing="ij")
    return x.reshape(-1), y.reshape(-1), z.reshape(-1)


def tg_analytic(cfg, U, k, E):
    x, y, _ = interior_xyz(cfg)
    return (-U * np.cos(k * x) * np.sin(k * y) * E, U * np.sin(k * x) * np.cos(k * y) * E,
            np.zeros(x.size))


def abc_analytic(cfg, U, k, E):
    x, y, z = interior_xyz(cfg)
    return (U * (np.sin(k * z) + np.cos(k * y)) * E, U * (np.sin(k * x) + np.cos(k * z)) * E,
            U * (np.sin(k * y) + np.cos(k * x)) * E)


def decaying_steps(nu, coef, frac):
    """std::size_t(frac / (nu * coef)) (validation/DecayingFlows.hpp:132-133)."""
    return int(frac / (nu * coef))


def parabola_roots(y, u):
    """Least-squares u = a + b y + c y^2 and its roots (validation/poiseuille.cpp:44-74)."""
    S = np.zeros((3, 4))
    for yk, uk in zip(y, u):
        p = (1.0, yk, yk * yk)
        for i in range(3):
            for j in range(3):
                S[i][j] += p[i] * p[j]
            S[i][3] += p[i] * uk
    for c in range(3):
        piv = c
        for r in range(c, 3):
            if abs(S[r][c]) > abs(S[piv][c]):
                piv = r
        S[[c, piv]] = S[[piv, c]]
        d = S[c][c]
        S[c] /= d
        for r in range(3):
            if r != c:
                S[r] -= S[r][c] * S[c]
    a, b, c2 = S[0][3], S[1][3], S[2][3]
    disc = math.sqrt(b * b - 4 * c2 * a)
    lo = (-b + disc) / (2 * c2)
    hi = (-b - disc) / (2 * c2)
    return (lo, hi) if lo <= hi else (hi, lo)


def poiseuille_config(op, tau, H, umax=0.01, dtype=abi.F64, storage=abi.SHIFTED):
    """validation/poiseuille.cpp:77-95: D2Q9, nx = 8, ny = H + 2, periodic x,
    walls (Solid rows y = 0 and y = H + 1), Guo force G along x."""
    nu = (tau - 0.5) * (1.0 / 3.0)
    omega = 1.0 / tau
    G = 8.0 * 1.0 * nu * umax / (H * H)
    wm = omega
    if op == abi.OP_TRT:
        wm = 1.0 / ((3.0 / 16.0) / (1.0 / omega - 0.5) + 0.5)
    cfg = abi.make_config(abi.D2Q9, op, storage, dtype, (8, H + 2, 1),
                          periodic=(True, False, True), omega=omega, omega_minus=wm,
                          force=(G, 0.0, 0.0))
    flags = np.zeros((H + 2, 8), dtype=np.uint8)
    flags[0, :] = abi.SOLID
    flags[H + 1, :] = abi.SOLID
    return cfg, flags.reshape(-1), nu, G


def poiseuille_measure(cfg, H, nu, G, rho, ux):
    """L2 against the analytic profile, mean density and wall slip
    (validation/poiseuille.cpp:110-140)."""
    d = abi.domain_of(cfg)
    sx, hy = d[6], d[4]
    cols = [((j + hy) * sx + 4) for j in range(1, H + 1)]
    yy = [float(j) for j in range(1, H + 1)]
    uu = [float(ux[c]) for c in cols]
    rho_mean = sum(float(rho[c]) for c in cols) / H
    A = G / (2.0 * rho_mean * nu)
    num = den = 0.0
    for j, u in zip(yy, uu):
        ana = A * (j - 0.5) * (H + 0.5 - j)
        num += (u - ana) ** 2
        den += ana * ana
    lo, _ = parabola_roots(yy, uu)
    return math.sqrt(num / den), rho_mean, 0.5 - lo


def poiseuille_run(solver, cfg, H, nu, G, max_steps=600000, probe_every=200):
    """March to steady state as validation/poiseuille.cpp:97-108 does, then
    measure.  `solver` has step(n) and compute_macroscopic()."""
    d = abi.domain_of(cfg)
    centre = ((cfg.ny // 2) + d[4]) * d[6] + cfg.nx // 2
    prev = 0.0
    taken = 0
    t = 0
    while t < max_steps:
        solver.step(probe_every)
        taken += probe_every
        rho, ux, _, _ = solver.compute_macroscopic()
        cur = float(ux[centre])
        if t > 0 and abs(cur - prev) < 1e-15 * abs(cur):
            break
        prev = cur
        t += probe_every
    rho, ux, _, _ = solver.compute_macroscopic()
    l2, rm, slip = poiseuille_measure(cfg, H, nu, G, rho, ux)
    return l2, rm, slip, taken


def taylor_green_box(cfg, u0):
    """Host copy of m3lb_initialize_taylor_green: the tgv3d_bench field
    (validation/tgv3d_bench.cpp:79-90) with one wavenumber per axis,
    k_a = 2 pi / n_a (for a cube, tgv3d_benchmark(cfg, u0))."""
    x, y, z, d = padded_coords(cfg)
    X = (2.0 * math.pi / cfg.nx) * (x - d[3])
    Y = (2.0 * math.pi / cfg.ny) * (y - d[4])
    Z = (2.0 * math.pi / cfg.nz) * (z - d[5])
    dr = 3.0 * u0 * u0 / 16.0
    rho = 1.0 + dr * (np.cos(2.0 * X) + np.cos(2.0 * Y)) * (np.cos(2.0 * Z) + 2.0)
    ux = u0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    uy = -u0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    return rho, ux, uy, np.zeros_like(ux)


def channel_flags(shape):
    """Interior flags of the bounce-back channel: rows y = 0 and y = ny - 1
    Solid (validation/poiseuille.cpp:86-90), x fastest."""
    nx, ny, nz = shape
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    fl[:, 0, :] = abi.SOLID
    fl[:, ny - 1, :] = abi.SOLID
    return fl.reshape(-1)


def serving_workload(n, seed=4, rate=40.0, lattice=abi.D3Q27, u0=0.04, edges=(32, 192)):
    """The batched-serving workload: n independent requests arriving as a
    Poisson stream of `rate` per second (seeded).  Half are periodic
    Taylor-Green boxes (amplitude u0, m3lb_initialize_taylor_green), half
    bounce-back channels (walls at y = 0 and ny - 1, periodic x and z, Guo
    force g_x ~ U[1e-7, 1e-5], starting at rest).  Each edge is drawn
    separately, log-uniform on [edges[0], edges[1]) (default [32, 192));
    tau ~ U[0.51, 0.8]; steps ~ U{200..3000}.
    Returns a list of (arrival offset in seconds, abi.Request)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        kind = rng.random() < 0.5
        lo, hi = edges
        e = [int(lo * (hi / lo) ** rng.random()) for _ in range(3)]
        if lattice == abi.D2Q9:
            e[2] = 1
        tau = rng.uniform(0.51, 0.8)
        steps = int(rng.integers(200, 3001))
        g = rng.uniform(1e-7, 1e-5)
        if kind:
            r = abi.make_request(e, steps, omega=1.0 / tau, init=abi.INIT_TAYLOR_GREEN,
                                 u=(u0, 0.0, 0.0))
        else:
            r = abi.make_request(e, steps, omega=1.0 / tau, periodic=(True, False, True),
                                 walls=abi.WALLS_Y, force=(g, 0.0, 0.0))
        out.append((t, r))
    return out


def sphere_channel(shape, n_spheres=200, rmin=6.0, rmax=14.0, seed=2, max_tries=200000):
    """Channel with spheres: Solid layers y in {0, ny-1} and z in {0, nz-1},
    plus n_spheres spheres with radii r ~ U[rmin, rmax] whose centres are
    drawn by seeded rejection sampling, accepted when the sphere keeps one
    free layer to every wall and to every accepted sphere.  A node is in a
    sphere when |x - c| <= r.  Returns (interior flags x fastest, sphere mask
    x fastest, list of (cx, cy, cz, r))."""
    nx, ny, nz = shape
    rng = np.random.Generator(np.random.MT19937(seed))
    spheres = []
    tries = 0
    while len(spheres) < n_spheres and tries < max_tries:
        tries += 1
        r = rng.uniform(rmin, rmax)
        c = (rng.uniform(0, nx), rng.uniform(r + 2.0, ny - 3.0 - r), rng.uniform(r + 2.0, nz - 3.0 - r))
        if c[1] > ny - 3.0 - r or c[2] > nz - 3.0 - r:
            continue
        ok = True
        for (a, b, e, q) in spheres:
            dx = abs(c[0] - a)
            dx = min(dx, nx - dx)  # periodic in x
            if dx * dx + (c[1] - b) ** 2 + (c[2] - e) ** 2 < (r + q + 1.0) ** 2:
                ok = False
                break
        if ok:
            spheres.append((c[0], c[1], c[2], r))
    body = np.zeros((nz, ny, nx), dtype=bool)
    for (a, b, e, q) in spheres:
        z0, z1 = int(math.floor(e - q)), int(math.ceil(e + q))
        y0, y1 = int(math.floor(b - q)), int(math.ceil(b + q))
        zz, yy = np.meshgrid(np.arange(z0, z1 + 1), np.arange(y0, y1 + 1), indexing="ij")
        xs = np.arange(int(math.floor(a - q)), int(math.ceil(a + q)) + 1)
        d2 = ((xs[None, None, :] - a) ** 2 + (yy[:, :, None] - b) ** 2 + (zz[:, :, None] - e) ** 2)
        inside = d2 <= q * q
        kz, ky, kx = np.nonzero(inside)
        body[zz[kz, ky], yy[kz, ky], xs[kx] % nx] = True
    flags = np.zeros((nz, ny, nx), dtype=np.uint8)
    flags[:, 0, :] = abi.SOLID
    flags[:, ny - 1, :] = abi.SOLID
    flags[0, :, :] = abi.SOLID
    flags[nz - 1, :, :] = abi.SOLID
    flags[body] = abi.SOLID
    return flags.reshape(-1), body.reshape(-1).astype(np.uint8), spheres


def shear_layer(cfg, u0, kappa=80.0, delta=0.05):
    """validation/shear_layer.cpp:43-54: double shear layer on an L x L
    periodic box, rho = 1, ux = u0 tanh(kappa (y/L - 1/4)) for y/L <= 1/2 and
    u0 tanh(kappa (3/4 - y/L)) above, uy = u0 delta sin(2 pi (x/L + 1/4))."""
    x, y, _, d = padded_coords(cfg)
    L = cfg.nx
    xx = (x - d[3]) / L
    yy = (y - d[4]) / L
    ux = np.where(yy <= 0.5, u0 * np.tanh(kappa * (yy - 0.25)), u0 * np.tanh(kappa * (0.75 - yy)))
    uy = u0 * delta * np.sin(2.0 * math.pi * (xx + 0.25))
    return np.ones_like(ux), ux, uy, np.zeros_like(ux)


# Outward wall normals of regularised velocity walls (boundary/Regularized.hpp
# NormalCode): 0 none, +x, -x, +y, -y, +z, -z, corner.
NRM_NONE, NRM_XP, NRM_XM, NRM_YP, NRM_YM, NRM_ZP, NRM_ZM, NRM_CORNER = range(8)


def cavity_walls(N, U, nz=1):
    """validation/cavity_cm.cpp:93-101: an N x N (x nz) box whose edge nodes
    are regularised walls -- the lid y = N-1 moving at (U, 0, 0), the other
    sides at rest, the four corner columns marked corner.  Returns interior
    (normals, tags) x fastest and the velocity table (ux, uy, uz, rho)."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(N), np.arange(N), indexing="ij")
    xl, xr, yb, yt = x == 0, x == N - 1, y == 0, y == N - 1
    nrm = np.zeros(x.shape, dtype=np.uint8)
    tag = np.zeros(x.shape, dtype=np.uint16)
    nrm[xr] = NRM_XP
    nrm[xl] = NRM_XM
    nrm[yb] = NRM_YM
    nrm[yt] = NRM_YP
    tag[yt] = 1
    nrm[(xl | xr) & (yb | yt)] = NRM_CORNER
    tag[(xl | xr) & (yb | yt)] = 0
    table = np.array([[0.0, 0.0, 0.0, 0.0], [U, 0.0, 0.0, 0.0]])
    return nrm.reshape(-1), tag.reshape(-1), table


def channel_walls(shape, u_top=0.0):
    """Regularised walls on the interior rows y = 0 (at rest) and y = ny - 1
    (moving at (u_top, 0, 0)): plane Couette or, with a force, Poiseuille."""
    nx, ny, nz = shape
    nrm = np.zeros((nz, ny, nx), dtype=np.uint8)
    tag = np.zeros((nz, ny, nx), dtype=np.uint16)
    nrm[:, 0, :] = NRM_YM
    nrm[:, ny - 1, :] = NRM_YP
    tag[:, ny - 1, :] = 1
    table = np.array([[0.0, 0.0, 0.0, 0.0], [u_top, 0.0, 0.0, 0.0]])
    return nrm.reshape(-1), tag.reshape(-1), table


def box_walls(shape, U):
    """A closed 3-D box: every face node a regularised wall (the face y = ny-1
    moving at (U, 0, 0), the others at rest), every node on an edge or vertex
    a corner at rest.  Returns interior (normals, tags) x fastest and the
    velocity table."""
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    faces = [(x == nx - 1, NRM_XP), (x == 0, NRM_XM), (y == ny - 1, NRM_YP), (y == 0, NRM_YM),
             (z == nz - 1, NRM_ZP), (z == 0, NRM_ZM)]
    nrm = np.zeros(x.shape, dtype=np.uint8)
    count = np.zeros(x.shape, dtype=np.int32)
    for m, code in faces:
        nrm[m] = code
        count += m
    tag = np.zeros(x.shape, dtype=np.uint16)
    tag[(y == ny - 1) & (count == 1)] = 1
    nrm[count > 1] = NRM_CORNER
    table = np.array([[0.0, 0.0, 0.0, 0.0], [U, 0.0, 0.0, 0.0]])
    return nrm.reshape(-1), tag.reshape(-1), table

NRM_OUT_XP, NRM_OUT_FREE, NRM_OUT_EQ = 8, 9, 10
SPEC_XP, SPEC_XM, SPEC_YP, SPEC_YM, SPEC_ZP, SPEC_ZM = 1, 2, 4, 8, 16, 32
