"""Continuous-batching solve server (gsg_server_* of include/gsg.h).

Requests are admitted in arrival order, first fit, while their device footprint (the
state, and the grid's operator tables for the first solve of a grid) fits the HBM
budget; a request that does not fit is passed by later ones for at most 8 admission
passes.  Resident solves advance together, `steps_per_tick` RK4 steps per scheduling
tick, and leave as soon as they finish, freeing room for queued requests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

STATUS = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class Server:
    def __init__(self, precision="fp64", hbm_budget_bytes=0, steps_per_tick=16, max_resident=0):
        cfg = L.ServerConfig()
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.precision = L.PREC[precision]
        cfg.steps_per_tick = int(steps_per_tick)
        cfg.max_resident = int(max_resident)
        self.precision = precision
        self._h = L.vp()
        L.check(L.lib().gsg_server_start(C.byref(cfg), C.byref(self._h)))
        self._lens = {}

    def submit(self, kind, D, k, n, y0, dt, nsteps, a=None, scheme="sparse"):
        desc = L.make_desc(D, k, n, kind, scheme, self.precision, coef=a)
        y = L.f64(y0)
        rid = C.c_int64()
        L.check(L.lib().gsg_server_submit(self._h, C.byref(desc), L.dptr(y), y.size, float(dt), int(nsteps),
                                          C.byref(rid)))
        self._lens[rid.value] = y.size
        return rid.value

    def poll(self, rid):
        st = C.c_int32()
        L.check(L.lib().gsg_server_poll(self._h, int(rid), C.byref(st)))
        return STATUS[st.value]

    def wait(self, rid):
        """Block until done; returns (final state, latency in seconds)."""
        y = np.zeros(self._lens.pop(rid))
        lat = C.c_int64()
        L.check(L.lib().gsg_server_wait(self._h, int(rid), L.dptr(y), y.size, C.byref(lat)))
        return y, lat.value * 1e-9

    def stats(self):
        s = L.ServerStats()
        L.check(L.lib().gsg_server_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def stop(self):
        if self._h:
            L.lib().gsg_server_stop(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
