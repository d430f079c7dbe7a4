"""ctypes mirror of include/m3lb.h: enums and the m3lb_config struct.

Pure Python, no device access: importing this module never loads a shared
library.  The numeric values are the header's; the header cites the reference
policy each one stands for.
"""
import ctypes

D2Q9 = 9
D3Q27 = 27

OP_NULL = 0
OP_BGK = 1
OP_BGK_PRODUCT = 2
OP_TRT = 3
OP_MRT = 4
OP_CM = 5
OP_TRT_PRODUCT = 6

RAW = 0
SHIFTED = 1

F64 = 0
F32 = 1
F16 = 2
BF16 = 3
I16 = 4
E2M13 = 5

FUSION_AUTO = 0
FUSION_OFF = 1
FUSION_ON = 2

FORCE_NONE = 0
FORCE_GUO = 1
FORCE_FIELD = 2

FLUID = 0
SOLID = 1
EXCLUDED = 2

OK = 0
EINVAL = -1
ENOMEM = -2
EDEVICE = -3
ESTATE = -4

OP_NAMES = {OP_NULL: "null", OP_BGK: "bgk", OP_BGK_PRODUCT: "bgk_product",
            OP_TRT: "trt", OP_MRT: "mrt", OP_CM: "cm", OP_TRT_PRODUCT: "trt_product"}
DTYPE_NAMES = {F64: "f64", F32: "f32", F16: "f16", BF16: "bf16", I16: "i16", E2M13: "e2m13"}


class Config(ctypes.Structure):
    """m3lb_config (include/m3lb.h)."""

    _fields_ = [
        ("lattice", ctypes.c_int32),
        ("op", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("omega", ctypes.c_double),
        ("omega_minus", ctypes.c_double),
        ("omega_bulk", ctypes.c_double),
        ("forcing", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("force", ctypes.c_double * 3),
    ]


def make_config(lattice=D3Q27, op=OP_CM, storage=SHIFTED, dtype=F64, shape=(8, 8, 8),
                periodic=(True, True, True), omega=1.0, omega_minus=1.0,
                omega_bulk=-1.0, force=None):
    """Builds a Config; force=None means NoForcing, a 3-tuple means uniform
    Guo, "field" a per-node Guo force (set_force_field)."""
    nx, ny, nz = shape
    c = Config()
    c.lattice = lattice
    c.op = op
    c.storage = storage
    c.dtype = dtype
    c.nx, c.ny, c.nz = int(nx), int(ny), int(nz)
    for a in range(3):
        c.periodic[a] = 1 if periodic[a] else 0
    c.omega = float(omega)
    c.omega_minus = float(omega_minus)
    c.omega_bulk = float(omega_bulk)
    if force is None:
        c.forcing = FORCE_NONE
        f = (0.0, 0.0, 0.0)
    elif isinstance(force, str) and force == "field":
        c.forcing = FORCE_FIELD
        f = (0.0, 0.0, 0.0)
    else:
        c.forcing = FORCE_GUO
        f = force
    for a in range(3):
        c.force[a] = float(f[a])
    c.reserved = 0
    return c


def domain_of(cfg):
    """Padded extents as src/grid/Domain.hpp:47-56 computes them (mpi_halo = 0).

    Returns (nx, ny, nz, hx, hy, hz, sx, sy, sz, n_padded)."""
    n = (cfg.nx, cfg.ny, cfg.nz)
    h = []
    for a in range(3):
        if n[a] <= 1 or cfg.periodic[a]:
            h.append(0)
        else:
            h.append(1)
    s = [n[a] + 2 * h[a] for a in range(3)]
    return (n[0], n[1], n[2], h[0], h[1], h[2], s[0], s[1], s[2], s[0] * s[1] * s[2])


# ---------------------------------------------------------------------------
# Batched serving (m3lb_srv_*).  No counterpart in the reference.
# ---------------------------------------------------------------------------
SHAPE_ROWS = 0
SHAPE_LINEAR = 1
SHAPE_AUTO = 2

TILE_BLOCK_RANGES = -2

INIT_UNIFORM = 0
INIT_TAYLOR_GREEN = 1

WALLS_NONE = 0
WALLS_Y = 1

REQ_QUEUED = 0
REQ_RESIDENT = 1
REQ_DONE = 2
REQ_CANCELLED = 3


class ServerConfig(ctypes.Structure):
    """m3lb_server_config (include/m3lb.h)."""

    _fields_ = [
        ("lattice", ctypes.c_int32),
        ("op", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("forcing", ctypes.c_int32),
        ("max_boxes", ctypes.c_int32),
        ("steps_per_chunk", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("tile_run", ctypes.c_int32),
        ("tile_shape", ctypes.c_int32),
        ("max_bypass", ctypes.c_int32),
        ("budget_bytes", ctypes.c_uint64),
        ("max_resident_nodes", ctypes.c_uint64),
    ]


class Request(ctypes.Structure):
    """m3lb_request (include/m3lb.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3),
        ("init", ctypes.c_int32),
        ("walls", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("keep", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("omega_minus", ctypes.c_double),
        ("omega_bulk", ctypes.c_double),
        ("force", ctypes.c_double * 3),
        ("rho0", ctypes.c_double),
        ("u", ctypes.c_double * 3),
        ("arrival", ctypes.c_double),
    ]


class Result(ctypes.Structure):
    """m3lb_result (include/m3lb.h)."""

    _fields_ = [
        ("id", ctypes.c_int64),
        ("state", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("arrival", ctypes.c_double),
        ("admitted", ctypes.c_double),
        ("completed", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("momentum", ctypes.c_double * 3),
    ]


class ServerStats(ctypes.Structure):
    """m3lb_server_stats (include/m3lb.h)."""

    _fields_ = [
        ("submitted", ctypes.c_int64),
        ("rejected", ctypes.c_int64),
        ("queued", ctypes.c_int64),
        ("resident", ctypes.c_int64),
        ("done", ctypes.c_int64),
        ("cancelled", ctypes.c_int64),
        ("chunks", ctypes.c_int64),
        ("steps", ctypes.c_int64),
        ("lattice_updates", ctypes.c_int64),
        ("bytes_used", ctypes.c_uint64),
        ("bytes_peak", ctypes.c_uint64),
        ("bytes_budget", ctypes.c_uint64),
        ("device_seconds", ctypes.c_double),
        ("grid", ctypes.c_int32),
        ("tile_run", ctypes.c_int32),
        ("tile_shape", ctypes.c_int32),
        ("steps_per_chunk", ctypes.c_int32),
        ("resident_nodes", ctypes.c_uint64),
    ]


class TuneResult(ctypes.Structure):
    """m3lb_tune_result (include/m3lb.h)."""

    _fields_ = [
        ("tile_shape", ctypes.c_int32),
        ("tile_run", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("candidates", ctypes.c_int32),
        ("best_glups", ctypes.c_double),
        ("initial_glups", ctypes.c_double),
    ]


def make_server_config(lattice=D3Q27, op=OP_CM, storage=SHIFTED, dtype=F32, forcing=FORCE_GUO,
                       budget_bytes=1 << 30, max_boxes=4096, steps_per_chunk=8, use_graph=True,
                       blocks_per_cu=8, tile_run=64, tile_shape=SHAPE_AUTO, max_bypass=64,
                       max_resident_nodes=0):
    c = ServerConfig()
    c.lattice, c.op, c.storage, c.dtype, c.forcing = lattice, op, storage, dtype, forcing
    c.max_boxes = int(max_boxes)
    c.steps_per_chunk = int(steps_per_chunk)
    c.use_graph = 1 if use_graph else 0
    c.blocks_per_cu = int(blocks_per_cu)
    c.tile_run = int(tile_run)
    c.tile_shape = int(tile_shape)
    c.max_bypass = int(max_bypass)
    c.budget_bytes = int(budget_bytes)
    c.max_resident_nodes = int(max_resident_nodes)
    return c


def make_request(shape, steps, omega=1.6, omega_minus=None, omega_bulk=-1.0,
                 periodic=(True, True, True), init=INIT_UNIFORM, walls=WALLS_NONE,
                 force=(0.0, 0.0, 0.0), rho0=1.0, u=(0.0, 0.0, 0.0), keep=False,
                 arrival=-1.0):
    r = Request()
    r.nx, r.ny, r.nz = (int(v) for v in shape)
    for a in range(3):
        r.periodic[a] = 1 if periodic[a] else 0
    r.init = init
    r.walls = walls
    r.steps = int(steps)
    r.keep = 1 if keep else 0
    r.omega = float(omega)
    r.omega_minus = float(omega if omega_minus is None else omega_minus)
    r.omega_bulk = float(omega_bulk)
    for a in range(3):
        r.force[a] = float(force[a])
        r.u[a] = float(u[a])
    r.rho0 = float(rho0)
    r.arrival = float(arrival)
    return r


def request_config(scfg, req):
    """The single-box m3lb_config that computes what `req` computes on a server
    configured by `scfg` (same kernel configuration and forcing policy)."""
    c = make_config(scfg.lattice, scfg.op, scfg.storage, scfg.dtype, (req.nx, req.ny, req.nz),
                    periodic=tuple(req.periodic), omega=req.omega, omega_minus=req.omega_minus,
                    omega_bulk=req.omega_bulk,
                    force=tuple(req.force) if scfg.forcing == FORCE_GUO else None)
    return c

CACHE_AUTO, CACHE_STREAM, CACHE_PLAIN = 0, 1, 2
