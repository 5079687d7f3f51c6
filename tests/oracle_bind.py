"""ctypes binding of the CPU oracle (oracle/_build/liborc.so).

Test infrastructure: only tests/, the smoke entry and bench.py's CPU
baseline use it, always as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "_build" / "liborc.so"
P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


class Oracle:
    def __init__(self, path: Path = LIB):
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = C.CDLL(str(path))
        self.lib.orc_sap_pairs.restype = C.c_size_t

    # ---- SAP ----
    def box_bounds(self, pos, quat, half, offset):
        n = len(pos)
        b = np.zeros(6 * n, np.uint32)
        rc = self.lib.orc_box_bounds(C.c_size_t(n), P(np.ascontiguousarray(pos, np.float32)),
                                     P(np.ascontiguousarray(quat, np.float32)),
                                     P(np.ascontiguousarray(half, np.float32)),
                                     P(np.ascontiguousarray(offset, np.float32)), P(b))
        if rc != 0:
            raise OverflowError("non-finite box bounds")
        return b.reshape(6, n)

    def sort_endpoints(self, bounds):
        n = bounds.shape[1]
        b = np.ascontiguousarray(bounds, np.uint32)
        vals = np.zeros((3, 2 * n), np.uint32)
        pays = np.zeros((3, 2 * n), np.uint32)
        for a in range(3):
            v = np.zeros(2 * n, np.uint32)
            p = np.zeros(2 * n, np.uint32)
            self.lib.orc_sort_endpoints(C.c_size_t(n), P(b), C.c_uint(a), P(v), P(p))
            vals[a], pays[a] = v, p
        return vals, pays

    def sap_pairs(self, bounds, shape_id=None, body_id=None, dynamic=None, plane_ids=()):
        n = bounds.shape[1]
        shape_id = np.arange(n, dtype=np.uint32) if shape_id is None else np.ascontiguousarray(shape_id, np.uint32)
        body_id = shape_id.copy() if body_id is None else np.ascontiguousarray(body_id, np.uint32)
        dynamic = np.ones(n, np.uint8) if dynamic is None else np.ascontiguousarray(dynamic, np.uint8)
        planes = np.ascontiguousarray(plane_ids, np.uint32)
        cap = 64 * n + len(planes) * n + 16
        out = np.zeros(2 * cap, np.uint32)
        m = self.lib.orc_sap_pairs(C.c_size_t(n), P(np.ascontiguousarray(bounds, np.uint32)), P(shape_id),
                                   P(body_id), P(dynamic), C.c_size_t(len(planes)),
                                   P(planes) if len(planes) else None, P(out), C.c_size_t(cap))
        assert m <= cap
        return out[: 2 * m].reshape(-1, 2)

    def found_lost(self, old, new):
        old = np.ascontiguousarray(old, np.uint32).reshape(-1, 2)
        new = np.ascontiguousarray(new, np.uint32).reshape(-1, 2)
        f = np.zeros(2 * len(new) + 2, np.uint32)
        l = np.zeros(2 * len(old) + 2, np.uint32)
        nf = C.c_size_t()
        nl = C.c_size_t()
        self.lib.orc_found_lost(C.c_size_t(len(old)), P(old), C.c_size_t(len(new)), P(new), P(f), C.byref(nf),
                                P(l), C.byref(nl))
        return f[: 2 * nf.value].reshape(-1, 2), l[: 2 * nl.value].reshape(-1, 2)
