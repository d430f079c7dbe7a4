"""Python mirror of the batched server (m3lb_srv_* in include/m3lb.h).

A Server keeps many independent boxes resident in one device arena and
advances them together; requests are submitted as abi.Request structs, results
come back as dicts.  Every call goes through the C-ABI; there is no CPU path.
The reference has no serving layer.
"""
import ctypes

import numpy as np

from . import abi
from .solver import _check, _d, _vp, load


def _result_dict(r):
    return {"id": int(r.id), "state": int(r.state), "steps": int(r.steps),
            "arrival": r.arrival, "admitted": r.admitted, "completed": r.completed,
            "mass": r.mass, "momentum": tuple(r.momentum)}


def footprint(scfg, req):
    """m3lb_srv_footprint: arena bytes of one request (no device call)."""
    L = load()
    out = ctypes.c_uint64()
    _check(L.m3lb_srv_footprint(ctypes.byref(scfg), ctypes.byref(req), ctypes.byref(out)))
    return out.value


class Server:
    def __init__(self, scfg):
        self.L = load()
        self.cfg = scfg
        self.Q = 9 if scfg.lattice == abi.D2Q9 else 27
        h = _vp()
        _check(self.L.m3lb_srv_create(ctypes.byref(scfg), ctypes.byref(h)))
        self.h = h
        self._req = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.m3lb_srv_destroy(self.h)
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

    def submit(self, req, flags=None):
        """Queues a request; flags = interior flags (x fastest) or None."""
        rid = ctypes.c_int64()
        fp = None
        if flags is not None:
            f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
            assert f.size == req.nx * req.ny * req.nz
            fp = f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        _check(self.L.m3lb_srv_submit(self.h, ctypes.byref(req), fp, ctypes.byref(rid)))
        self._req[rid.value] = req
        return rid.value

    def progress(self, wait_us=0):
        n = ctypes.c_int32()
        _check(self.L.m3lb_srv_progress(self.h, int(wait_us), ctypes.byref(n)))
        return n.value

    def drain(self):
        _check(self.L.m3lb_srv_drain(self.h))

    def poll(self, max_results=4096):
        buf = (abi.Result * max_results)()
        n = ctypes.c_int32()
        _check(self.L.m3lb_srv_poll(self.h, buf, max_results, ctypes.byref(n)))
        return [_result_dict(buf[k]) for k in range(n.value)]

    def result(self, rid):
        r = abi.Result()
        _check(self.L.m3lb_srv_result(self.h, int(rid), ctypes.byref(r)))
        return _result_dict(r)

    def _np(self, rid):
        req = self._req[rid]
        cfg = abi.request_config(self.cfg, req)
        return abi.domain_of(cfg)[9]

    def fetch_macroscopic(self, rid):
        n = self._np(rid)
        out = [np.zeros(n) for _ in range(4)]
        _check(self.L.m3lb_srv_fetch_macroscopic(self.h, int(rid), *[_d(a) for a in out]))
        return out

    def fetch_populations(self, rid):
        out = np.zeros(self._np(rid) * self.Q)
        _check(self.L.m3lb_srv_fetch_populations(self.h, int(rid), _d(out)))
        return out.reshape(-1, self.Q)

    def release(self, rid):
        _check(self.L.m3lb_srv_release(self.h, int(rid)))

    def cancel(self, rid):
        _check(self.L.m3lb_srv_cancel(self.h, int(rid)))

    def stats(self):
        s = abi.ServerStats()
        _check(self.L.m3lb_srv_stats(self.h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in abi.ServerStats._fields_}

    def clock(self):
        t = ctypes.c_double()
        _check(self.L.m3lb_srv_clock(self.h, ctypes.byref(t)))
        return t.value

    def autotune(self, steps=5):
        r = abi.TuneResult()
        _check(self.L.m3lb_srv_autotune(self.h, int(steps), ctypes.byref(r)))
        return {name: getattr(r, name) for name, _ in abi.TuneResult._fields_}
