"""The solver-order authority (host C++ in libfyzz_hip.so): the native
scene's island graphs kept from each step's broadphase and touch changes,
and the solver order they give (include/fyzz_hip.h, fyzz_hip_islands_*).
It needs no device: the graphs live on the host, as the reference keeps
them (scene.cpp:5119-5952, 6148-6262, 6465-6536, 7174-7260)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FyzzError, lib


class _Events(C.Structure):
    _fields_ = [("pairs_found", C.c_void_p), ("pairs_found_count", C.c_uint32),
                ("pairs_lost", C.c_void_p), ("pairs_lost_count", C.c_uint32),
                ("touch_found", C.c_void_p), ("touch_found_count", C.c_uint32),
                ("touch_lost", C.c_void_p), ("touch_lost_count", C.c_uint32),
                ("ready", C.c_void_p)]


class _Order(C.Structure):
    _fields_ = [("managers", C.c_void_p), ("manager_capacity", C.c_uint32), ("manager_count", C.c_uint32),
                ("task_ends", C.c_void_p), ("task_capacity", C.c_uint32), ("task_count", C.c_uint32),
                ("sleeping", C.c_void_p)]


def _pairs(p):
    a = np.ascontiguousarray(np.asarray(p, dtype=np.uint32).reshape(-1, 2))
    return a, a.ctypes.data if len(a) else None, len(a)


class IslandOrder:
    """dynamic[n]: True for a dynamic body, False for a static one."""

    def __init__(self, dynamic):
        d = np.ascontiguousarray(np.asarray(dynamic, dtype=np.uint8))
        self.n = len(d)
        h = C.c_void_p()
        rc = lib().fyzz_hip_islands_create(self.n, d.ctypes.data, C.byref(h))
        if rc != 0:
            raise FyzzError(rc, "fyzz_hip_islands_create failed")
        self.h = h
        self.held = 0  # broadphase pairs held: the solver order is at most these

    def close(self):
        if self.h:
            lib().fyzz_hip_islands_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, pairs_found, pairs_lost, touch_found, touch_lost, ready=None):
        """One step's changes (ascending (low, high) pairs); returns
        (managers [k, 2] in solver order, task ends, sleeping flags)."""
        keep = [_pairs(x) for x in (pairs_found, pairs_lost, touch_found, touch_lost)]
        r = None if ready is None else np.ascontiguousarray(np.asarray(ready, dtype=np.uint8))
        ev = _Events(keep[0][1], keep[0][2], keep[1][1], keep[1][2], keep[2][1], keep[2][2], keep[3][1], keep[3][2],
                     None if r is None else r.ctypes.data)
        held = self.held + keep[0][2] - keep[1][2]
        mg = np.zeros((max(held, 1), 2), np.uint32)
        te = np.zeros(self.n + 1, np.uint32)
        sl = np.zeros(self.n, np.uint8)
        out = _Order(mg.ctypes.data, len(mg), 0, te.ctypes.data, len(te), 0, sl.ctypes.data)
        rc = lib().fyzz_hip_islands_step(self.h, C.byref(ev), C.byref(out))
        if rc == 0:
            self.held = held
        else:
            msg = lib().fyzz_hip_islands_last_error(self.h)
            raise FyzzError(rc, msg.decode() if msg else "")
        return mg[:out.manager_count].copy(), te[:out.task_count].copy(), sl.astype(bool)
