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

RAW = 0
SHIFTED = 1

F64 = 0
F32 = 1
F16 = 2
BF16 = 3

FORCE_NONE = 0
FORCE_GUO = 1

FLUID = 0
SOLID = 1
EXCLUDED = 2

OK = 0
EINVAL = -1
ENOMEM = -2
EDEVICE = -3
ESTATE = -4

OP_NAMES = {OP_NULL: "null", OP_BGK: "bgk", OP_BGK_PRODUCT: "bgk_product",
            OP_TRT: "trt", OP_MRT: "mrt", OP_CM: "cm"}
DTYPE_NAMES = {F64: "f64", F32: "f32", F16: "f16", BF16: "bf16"}


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
    """Builds a Config; force=None means NoForcing, a 3-tuple means Guo."""
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
