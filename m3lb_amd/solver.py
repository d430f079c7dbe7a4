"""Python mirror of FluidSolver<L, EsotericPull<L>, Collision> over libm3lb_hip.so.

Method names follow the reference's public solver API
(src/solver/FluidSolver.hpp:42-637); every call goes through the C-ABI in
include/m3lb.h.  There is no CPU fallback: if the library is missing or a call
fails, a RuntimeError carries the library's message.
"""
import ctypes
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libm3lb_hip.so")
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_cfgp = ctypes.POINTER(abi.Config)
_vp = ctypes.c_void_p

# Every entry point of include/m3lb.h with its ctypes signature.
SIGNATURES = {
    "m3lb_domain": (ctypes.c_int, [_cfgp, _i64p]),
    "m3lb_create": (ctypes.c_int, [_cfgp, ctypes.POINTER(_vp)]),
    "m3lb_destroy": (ctypes.c_int, [_vp]),
    "m3lb_last_error": (ctypes.c_char_p, []),
    "m3lb_set_geometry": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint8)]),
    "m3lb_initialize": (ctypes.c_int, [_vp] + [ctypes.c_double] * 4),
    "m3lb_initialize_field": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "m3lb_step": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "m3lb_compute_macroscopic": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "m3lb_macroscopic_stored": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "m3lb_total_mass": (ctypes.c_int, [_vp, _dp]),
    "m3lb_total_momentum": (ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    "m3lb_storage_sum_exact": (ctypes.c_int, [_vp, _i64p]),
    "m3lb_energy_enstrophy": (ctypes.c_int, [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int32)]),
    "m3lb_gather_populations": (ctypes.c_int, [_vp, _dp]),
    "m3lb_seed_populations": (ctypes.c_int, [_vp, _dp]),
    "m3lb_storage": (ctypes.c_int, [_vp, _dp]),
    "m3lb_timestep": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "m3lb_set_force": (ctypes.c_int, [_vp] + [ctypes.c_double] * 3),
    "m3lb_set_force_field": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "m3lb_set_regularized_walls": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint8),
                                                  ctypes.POINTER(ctypes.c_uint16), _dp,
                                                  ctypes.c_int32]),
    "m3lb_set_fd_corners": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "m3lb_set_cache_policy": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "m3lb_set_force_correction": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "m3lb_set_outflow_order": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "m3lb_set_outflow_density": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m3lb_set_wall_velocity_scale": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m3lb_set_wall_density": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m3lb_set_wall_velocity_field": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "m3lb_set_wall_velocity_field_scale": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m3lb_set_specular_walls": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint8)]),
    "m3lb_set_specular_nodes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint8)]),
    "m3lb_set_omega": (ctypes.c_int, [_vp] + [ctypes.c_double] * 3),
    "m3lb_get_scales": (ctypes.c_int, [_vp, _dp]),
    "m3lb_set_scales": (ctypes.c_int, [_vp, _dp]),
    "m3lb_calibrate_scales": (ctypes.c_int, [_vp]),
    "m3lb_time_steps": (ctypes.c_int, [_vp, ctypes.c_int32, _dp]),
    "m3lb_device_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "m3lb_set_tile_order": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "m3lb_tile_order_map": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_uint32)]),
    "m3lb_stream_probe": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, _dp]),
    "m3lb_kernel_build_id": (ctypes.c_char_p, []),
    "m3lb_set_step_fusion": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_int32]),
    "m3lb_fused_passes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "m3lb_totals_static": (ctypes.c_int, [_vp, _dp]),
    "m3lb_test_fail_fused_waits": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "m3lb_initialize_taylor_green": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m3lb_momentum_exchange": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint8), _dp]),
    # batched server
    "m3lb_srv_create": (ctypes.c_int, [ctypes.POINTER(abi.ServerConfig), ctypes.POINTER(_vp)]),
    "m3lb_srv_destroy": (ctypes.c_int, [_vp]),
    "m3lb_srv_footprint": (ctypes.c_int, [ctypes.POINTER(abi.ServerConfig),
                                          ctypes.POINTER(abi.Request),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "m3lb_srv_submit": (ctypes.c_int, [_vp, ctypes.POINTER(abi.Request),
                                       ctypes.POINTER(ctypes.c_uint8), _i64p]),
    "m3lb_srv_progress": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "m3lb_srv_drain": (ctypes.c_int, [_vp]),
    "m3lb_srv_poll": (ctypes.c_int, [_vp, ctypes.POINTER(abi.Result), ctypes.c_int32,
                                     ctypes.POINTER(ctypes.c_int32)]),
    "m3lb_srv_result": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.POINTER(abi.Result)]),
    "m3lb_srv_fetch_macroscopic": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp, _dp, _dp]),
    "m3lb_srv_fetch_populations": (ctypes.c_int, [_vp, ctypes.c_int64, _dp]),
    "m3lb_srv_release": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "m3lb_srv_cancel": (ctypes.c_int, [_vp, ctypes.c_int64]),
    "m3lb_srv_stats": (ctypes.c_int, [_vp, ctypes.POINTER(abi.ServerStats)]),
    "m3lb_srv_clock": (ctypes.c_int, [_vp, _dp]),
    "m3lb_srv_autotune": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(abi.TuneResult)]),
}


def stream_probe(nbytes, nontemporal=True, reps=5):
    """m3lb_stream_probe: best milliseconds for an in-place stream of nbytes."""
    L = load()
    ms = ctypes.c_double()
    _check(L.m3lb_stream_probe(int(nbytes), 1 if nontemporal else 0, int(reps), ctypes.byref(ms)))
    return ms.value


def kernel_build_id():
    """m3lb_kernel_build_id: the hash of the sources the loaded library was built from."""
    return load().m3lb_kernel_build_id().decode()


def tile_order_map(grid, run):
    """m3lb_tile_order_map: the step kernel's block -> tile map."""
    L = load()
    out = np.zeros(int(grid), dtype=np.uint32)
    _check(L.m3lb_tile_order_map(int(grid), int(run),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
    return out


def load():
    """Loads libm3lb_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                           "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        msg = _LIB.m3lb_last_error().decode()
        raise RuntimeError(f"m3lb error {rc}: {msg}")


def _d(a):
    return a.ctypes.data_as(_dp)


def domain(cfg):
    """m3lb_domain: (nx, ny, nz, hx, hy, hz, sx, sy, sz, n_padded)."""
    L = load()
    out = (ctypes.c_int64 * 10)()
    _check(L.m3lb_domain(ctypes.byref(cfg), out))
    return tuple(out)


class FluidSolver:
    """One device-resident lattice (FluidSolver.hpp:34-69)."""

    def __init__(self, cfg):
        self.L = load()
        self.cfg = cfg
        self.Q = 9 if cfg.lattice == abi.D2Q9 else 27
        self.dom = domain(cfg)
        self.np = self.dom[9]
        h = _vp()
        _check(self.L.m3lb_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.m3lb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- setup ---------------------------------------------------------------
    def set_geometry(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        assert f.size == self.cfg.nx * self.cfg.ny * self.cfg.nz
        _check(self.L.m3lb_set_geometry(self.h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def initialize(self, rho0, ux=0.0, uy=0.0, uz=0.0):
        _check(self.L.m3lb_initialize(self.h, rho0, ux, uy, uz))

    def initialize_taylor_green(self, u0):
        """m3lb_initialize_taylor_green: the tgv3d_bench vortex evaluated on the
        device (cases.taylor_green_box gives the same field on the host)."""
        _check(self.L.m3lb_initialize_taylor_green(self.h, float(u0)))

    def initialize_field(self, rho, ux, uy, uz):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (rho, ux, uy, uz)]
        for a in arrs:
            assert a.size == self.np
        _check(self.L.m3lb_initialize_field(self.h, *[_d(a) for a in arrs]))

    def seed_populations(self, pops):
        a = np.ascontiguousarray(pops, dtype=np.float64).reshape(-1)
        assert a.size == self.np * self.Q
        _check(self.L.m3lb_seed_populations(self.h, _d(a)))

    # -- stepping ------------------------------------------------------------
    def step(self, n=1, store_macroscopic=False):
        _check(self.L.m3lb_step(self.h, int(n), 1 if store_macroscopic else 0))

    def time_steps(self, n):
        ms = ctypes.c_double()
        _check(self.L.m3lb_time_steps(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    # -- readout -------------------------------------------------------------
    def compute_macroscopic(self):
        out = [np.zeros(self.np) for _ in range(4)]
        _check(self.L.m3lb_compute_macroscopic(self.h, *[_d(a) for a in out]))
        return out

    def macroscopic_stored(self):
        out = [np.zeros(self.np) for _ in range(4)]
        _check(self.L.m3lb_macroscopic_stored(self.h, *[_d(a) for a in out]))
        return out

    def total_mass(self):
        v = ctypes.c_double()
        _check(self.L.m3lb_total_mass(self.h, ctypes.byref(v)))
        return v.value

    def total_momentum(self, axis):
        v = ctypes.c_double()
        _check(self.L.m3lb_total_momentum(self.h, int(axis), ctypes.byref(v)))
        return v.value

    def storage_sum_exact(self):
        o = (ctypes.c_int64 * 2)()
        _check(self.L.m3lb_storage_sum_exact(self.h, o))
        return (o[0], o[1])

    def energy_enstrophy(self):
        e, w, ok = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
        _check(self.L.m3lb_energy_enstrophy(self.h, ctypes.byref(e), ctypes.byref(w),
                                            ctypes.byref(ok)))
        return e.value, w.value, bool(ok.value)

    def momentum_exchange(self, body=None):
        """m3lb_momentum_exchange: (Fx, Fy, Fz) on the body (interior mask,
        x fastest; None = every Solid node)."""
        f = (ctypes.c_double * 3)()
        bp = None
        if body is not None:
            b = np.ascontiguousarray(body, dtype=np.uint8).reshape(-1)
            assert b.size == self.cfg.nx * self.cfg.ny * self.cfg.nz
            bp = b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        _check(self.L.m3lb_momentum_exchange(self.h, bp, f))
        return tuple(f)

    def gather_populations(self):
        out = np.zeros(self.np * self.Q)
        _check(self.L.m3lb_gather_populations(self.h, _d(out)))
        return out.reshape(self.np, self.Q)

    def storage(self):
        out = np.zeros(self.np * self.Q)
        _check(self.L.m3lb_storage(self.h, _d(out)))
        return out.reshape(self.np, self.Q)

    def timestep(self):
        t = ctypes.c_uint64()
        _check(self.L.m3lb_timestep(self.h, ctypes.byref(t)))
        return t.value

    def set_regularized_walls(self, normals, tags, table):
        """m3lb_set_regularized_walls: interior normal codes and table indices
        (x fastest) and the (ux, uy, uz, rho) table."""
        nrm = np.ascontiguousarray(normals, dtype=np.uint8).reshape(-1)
        tg = np.ascontiguousarray(tags, dtype=np.uint16).reshape(-1)
        tab = np.ascontiguousarray(table, dtype=np.float64).reshape(-1, 4)
        n = self.cfg.nx * self.cfg.ny * self.cfg.nz
        assert nrm.size == n and tg.size == n
        _check(self.L.m3lb_set_regularized_walls(
            self.h, nrm.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            tg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), _d(tab), int(tab.shape[0])))

    def set_cache_policy(self, policy):
        """abi.CACHE_AUTO / CACHE_STREAM / CACHE_PLAIN (results identical)."""
        _check(self.L.m3lb_set_cache_policy(self.h, int(policy)))

    def set_fd_corners(self, on):
        _check(self.L.m3lb_set_fd_corners(self.h, 1 if on else 0))

    def set_force_correction(self, on):
        _check(self.L.m3lb_set_force_correction(self.h, 1 if on else 0))

    def set_outflow_order(self, k):
        _check(self.L.m3lb_set_outflow_order(self.h, int(k)))

    def set_outflow_density(self, rho):
        _check(self.L.m3lb_set_outflow_density(self.h, float(rho)))

    def set_wall_velocity_scale(self, sc):
        _check(self.L.m3lb_set_wall_velocity_scale(self.h, float(sc)))

    def set_wall_density(self, rho):
        _check(self.L.m3lb_set_wall_density(self.h, float(rho)))

    def set_wall_velocity_field(self, ux, uy, uz):
        """One wall velocity per interior node, x fastest."""
        n = self.cfg.nx * self.cfg.ny * self.cfg.nz
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (ux, uy, uz)]
        assert all(a.size == n for a in arrs)
        _check(self.L.m3lb_set_wall_velocity_field(self.h, *[_d(a) for a in arrs]))

    def set_wall_velocity_field_scale(self, sc):
        _check(self.L.m3lb_set_wall_velocity_field_scale(self.h, float(sc)))

    def set_specular_walls(self, normals):
        a = np.ascontiguousarray(normals, dtype=np.uint8).reshape(-1)
        assert a.size == self.cfg.nx * self.cfg.ny * self.cfg.nz
        _check(self.L.m3lb_set_specular_walls(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def set_specular_nodes(self, faces):
        a = np.ascontiguousarray(faces, dtype=np.uint8).reshape(-1)
        assert a.size == self.cfg.nx * self.cfg.ny * self.cfg.nz
        _check(self.L.m3lb_set_specular_nodes(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def set_force_field(self, fx, fy, fz):
        """m3lb_set_force_field: per-node Guo force, padded arrays."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (fx, fy, fz)]
        for a in arrs:
            assert a.size == self.np
        _check(self.L.m3lb_set_force_field(self.h, *[_d(a) for a in arrs]))

    def set_force(self, fx, fy, fz):
        _check(self.L.m3lb_set_force(self.h, fx, fy, fz))

    def set_omega(self, omega, omega_minus=None, omega_bulk=-1.0):
        if omega_minus is None:
            omega_minus = omega
        _check(self.L.m3lb_set_omega(self.h, omega, omega_minus, omega_bulk))

    def scales(self):
        out = np.zeros(self.Q)
        _check(self.L.m3lb_get_scales(self.h, _d(out)))
        return out

    def set_scales(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64)
        _check(self.L.m3lb_set_scales(self.h, _d(a)))

    def calibrate_scales(self):
        _check(self.L.m3lb_calibrate_scales(self.h))

    def set_step_fusion(self, mode, stripes=0, lag=0, blocks_per_cu=0, group=0):
        """m3lb_set_step_fusion: pair steps in one pass (bit-identical results)."""
        _check(self.L.m3lb_set_step_fusion(self.h, int(mode), int(stripes), int(lag),
                                           int(blocks_per_cu), int(group)))

    def totals_static(self):
        """m3lb_totals_static: (mass, Mx, My, Mz) with no device allocation."""
        out = np.zeros(4)
        _check(self.L.m3lb_totals_static(self.h, _d(out)))
        return out

    def fused_passes(self):
        """m3lb_fused_passes: pairs of steps run as one pass so far."""
        n = ctypes.c_uint64()
        _check(self.L.m3lb_fused_passes(self.h, ctypes.byref(n)))
        return n.value

    def test_fail_fused_waits(self, on):
        """m3lb_test_fail_fused_waits: failure-path test hook."""
        _check(self.L.m3lb_test_fail_fused_waits(self.h, 1 if on else 0))

    def set_tile_order(self, run):
        _check(self.L.m3lb_set_tile_order(self.h, int(run)))

    def device_bytes(self):
        b = ctypes.c_uint64()
        _check(self.L.m3lb_device_bytes(self.h, ctypes.byref(b)))
        return b.value
