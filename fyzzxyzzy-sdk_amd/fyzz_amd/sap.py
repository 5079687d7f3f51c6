"""Python mirror of the GPU broadphase producer (GpuBroadphaseProducer,
reference src/native/scene.h:383-505) over the C-ABI.

Device buffers are torch tensors on the context's device (torch is only
the allocator and stream owner here); all compute runs in libfyzz_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Context, FyzzError, lib  # noqa: F401

vp = C.c_void_p


def _ptr(t):
    return vp(t.data_ptr()) if t is not None else None


def box_bounds(ctx: Context, pos, quat, half, offset):
    """Encoded bounds [6, n] (u32 as int32 tensor) of device arrays."""
    import torch
    n = pos.shape[0]
    out = torch.empty((6, n), dtype=torch.int32, device=pos.device)
    err = torch.zeros(1, dtype=torch.int32, device=pos.device)
    ctx.call("fyzz_hip_sap_box_bounds", C.c_uint32(n), _ptr(pos), _ptr(quat), _ptr(half), _ptr(offset),
             _ptr(out), _ptr(err))
    return out, err


def sort_endpoints(ctx: Context, bounds):
    """Sorted (values, payloads), each [3, 2n], from bounds [6, n]."""
    import torch
    n = bounds.shape[1]
    vals = torch.empty((3, 2 * n), dtype=torch.int32, device=bounds.device)
    pays = torch.empty((3, 2 * n), dtype=torch.int32, device=bounds.device)
    words = lib().fyzz_hip_sap_sort_scratch_words(n)
    scratch = torch.empty(int(words), dtype=torch.int32, device=bounds.device)
    ctx.call("fyzz_hip_sap_sort_endpoints", C.c_uint32(n), _ptr(bounds), _ptr(vals), _ptr(pays), _ptr(scratch),
             C.c_uint64(int(words)))
    return vals, pays


def sort_endpoints_fixed(ctx: Context, bounds):
    """As sort_endpoints, through the entry point that allocates nothing
    (its workspace is the library's fixed per-device array)."""
    import torch
    n = bounds.shape[1]
    vals = torch.empty((3, 2 * n), dtype=torch.int32, device=bounds.device)
    pays = torch.empty((3, 2 * n), dtype=torch.int32, device=bounds.device)
    ctx.call("fyzz_hip_sap_sort_endpoints_fixed", C.c_uint32(n), _ptr(bounds), _ptr(vals), _ptr(pays))
    return vals, pays


class SapBroadphase:
    """Incremental SAP over n boxes in slot order (increasing shape id)."""

    def __init__(self, ctx: Context, body_ids, dynamic, pair_capacity: int | None = None,
                 delta_capacity: int | None = None):
        n = len(body_ids)
        self.ctx = ctx
        self.n = n
        b = np.ascontiguousarray(body_ids, np.uint32)
        d = np.ascontiguousarray(dynamic, np.uint8)
        pc = pair_capacity if pair_capacity is not None else max(8 * n, 4096)
        dc = delta_capacity if delta_capacity is not None else max(8 * n, 4096)
        h = vp()
        ctx.call("fyzz_hip_sap_create", C.c_uint32(n), b.ctypes.data_as(vp), d.ctypes.data_as(vp),
                 C.c_uint32(pc), C.c_uint32(dc), C.byref(h))
        self.h = h
        ctx.adopt(self)

    def close(self):
        if self.h:
            lib().fyzz_hip_sap_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise FyzzError(rc, lib().fyzz_hip_last_error(self.ctx.h).decode())

    def advance(self, pos, quat, half, offset, dirty=None, force_full=False):
        nd = 0 if dirty is None else int(dirty.numel())
        self._check(lib().fyzz_hip_sap_advance(self.h, _ptr(pos), _ptr(quat), _ptr(half), _ptr(offset),
                                               C.c_uint32(nd), _ptr(dirty), 1 if force_full else 0))

    def counts(self):
        np_, nf, nl = C.c_uint64(), C.c_uint64(), C.c_uint64()
        fl = C.c_uint32()
        self._check(lib().fyzz_hip_sap_counts(self.h, C.byref(np_), C.byref(nf), C.byref(nl), C.byref(fl)))
        return np_.value, nf.value, nl.value, fl.value

    def fetch(self, which: int) -> np.ndarray:
        """0 live pairs, 1 found, 2 lost: sorted (lo, hi) slot pairs."""
        total = C.c_uint64()
        self._check(lib().fyzz_hip_sap_fetch(self.h, which, None, 0, C.byref(total)))
        out = np.zeros(2 * total.value + 2, np.uint32)
        self._check(lib().fyzz_hip_sap_fetch(self.h, which, out.ctypes.data_as(vp), C.c_uint64(total.value),
                                             C.byref(total)))
        return out[: 2 * total.value].reshape(-1, 2)

    def endpoints(self):
        import torch
        dev = torch.device("cuda", 0)
        vals = torch.empty((3, 2 * self.n), dtype=torch.int32, device=dev)
        pays = torch.empty((3, 2 * self.n), dtype=torch.int32, device=dev)
        self._check(lib().fyzz_hip_sap_endpoints(self.h, _ptr(vals), _ptr(pays)))
        return vals, pays
