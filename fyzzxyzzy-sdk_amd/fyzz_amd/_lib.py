"""ctypes binding of the C-ABI in include/fyzz_hip.h.

The library is loaded from this package directory (libfyzz_hip.so, built
in-tree by fyzzxyzzy-sdk_amd/build.py). There is no fallback: a missing or
unloadable library raises ImportError-derived LibraryMissing.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfyzz_hip.so"

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p

# name -> (restype, argtypes). Mirrors include/fyzz_hip.h one to one.
SIGNATURES = {
    "fyzz_hip_version": (C.c_char_p, []),
    "fyzz_hip_context_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "fyzz_hip_context_destroy": (C.c_int, [vp]),
    "fyzz_hip_last_error": (C.c_char_p, [vp]),
    "fyzz_hip_set_stream": (C.c_int, [vp, vp]),
    "fyzz_hip_synchronize": (C.c_int, [vp]),
    "fyzz_hip_set_cooperative": (C.c_int, [vp, C.c_int]),
    "fyzz_hip_profile_enable": (C.c_int, [vp, C.c_int]),
    "fyzz_hip_profile_read": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_double), u64p]),
    "fyzz_hip_sap_box_bounds": (C.c_int, [vp, C.c_uint32, vp, vp, vp, vp, vp, vp]),
    "fyzz_hip_sap_sort_endpoints": (C.c_int, [vp, C.c_uint32, vp, vp, vp, vp, C.c_uint64]),
    "fyzz_hip_sap_sort_endpoints_fixed": (C.c_int, [vp, C.c_uint32, vp, vp, vp]),
    "fyzz_hip_sap_sort_fixed_max_boxes": (C.c_uint32, []),
    "fyzz_hip_sap_sort_scratch_words": (C.c_uint64, [C.c_uint32]),
    "fyzz_hip_sap_create": (C.c_int, [vp, C.c_uint32, vp, vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]),
    "fyzz_hip_sap_destroy": (C.c_int, [vp]),
    "fyzz_hip_sap_advance": (C.c_int, [vp, vp, vp, vp, vp, C.c_uint32, vp, C.c_int]),
    "fyzz_hip_sap_counts": (C.c_int, [vp, u64p, u64p, u64p, u32p]),
    "fyzz_hip_sap_fetch": (C.c_int, [vp, C.c_int, vp, C.c_uint64, u64p]),
    "fyzz_hip_sap_endpoints": (C.c_int, [vp, vp, vp]),
    "fyzz_hip_pcm_cache_bytes": (C.c_uint32, []),
    "fyzz_hip_pcm_box_box": (C.c_int, [vp, C.c_uint32, vp, vp, vp, vp, vp, vp, C.c_float, vp, vp, vp, vp, vp]),
    "fyzz_hip_pcm_plane_box": (C.c_int, [vp, C.c_uint32, vp, vp, vp, vp, vp, vp, C.c_float, vp, vp, vp, vp]),
    "fyzz_hip_pcm_candidates": (C.c_int, [vp, C.c_uint32, vp, vp, C.c_float, vp, vp, vp, vp]),
    "fyzz_hip_pcm_gjk_epa": (C.c_int, [vp, C.c_uint32, vp, vp, C.c_float, vp, vp]),
    "fyzz_hip_dy_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]),
    "fyzz_hip_dy_destroy": (C.c_int, [vp]),
    "fyzz_hip_dy_partition": (C.c_int, [vp, C.c_uint32, vp, C.c_uint32, vp, vp]),
    "fyzz_hip_dy_prepare": (C.c_int, [vp, C.c_float, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, vp,
                                      C.c_uint32, vp]),
    "fyzz_hip_dy_solve": (C.c_int, [vp]),
    "fyzz_hip_dy_integrate": (C.c_int, [vp, vp, vp]),
    "fyzz_hip_dy_step": (C.c_int, [vp, C.c_float, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, vp, C.c_uint32,
                                   vp, vp, vp]),
    "fyzz_hip_dy_status": (C.c_int, [vp, u32p, u32p, u32p]),
    "fyzz_hip_dy_friction_anchors": (C.c_int, [vp, vp]),
    "fyzz_hip_dy_set_tape": (C.c_int, [vp, C.c_int]),
    "fyzz_hip_dy_adjoint": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "fyzz_hip_world_create": (C.c_int, [vp, C.c_uint32, C.POINTER(vp)]),
    "fyzz_hip_world_destroy": (C.c_int, [vp]),
    "fyzz_hip_world_add_static_floor": (C.c_int, [vp, C.c_float, u32p]),
    "fyzz_hip_world_add_dynamic_box": (C.c_int, [vp, vp, C.c_float, u32p]),
    "fyzz_hip_world_add_dynamic_boxes": (C.c_int, [vp, C.c_uint32, vp, C.c_float, vp, vp, vp, vp, C.c_float,
                                                   C.c_float, C.c_float, u32p]),
    "fyzz_hip_world_set_body_pose": (C.c_int, [vp, C.c_uint32, vp, vp]),
    "fyzz_hip_world_set_body_velocity": (C.c_int, [vp, C.c_uint32, vp, vp]),
    "fyzz_hip_world_set_body_material": (C.c_int, [vp, C.c_uint32, C.c_float, C.c_float, C.c_float]),
    "fyzz_hip_world_set_body_dynamics": (C.c_int, [vp, C.c_uint32, vp]),
    "fyzz_hip_world_set_scenes": (C.c_int, [vp, C.c_uint32, vp]),
    "fyzz_hip_world_set_materials": (C.c_int, [vp, C.c_uint32, C.c_uint32, vp, vp, vp]),
    "fyzz_hip_world_record_dy_inputs": (C.c_int, [vp, vp, vp, C.c_uint32, u32p]),
    "fyzz_hip_world_copy_bodies": (C.c_int, [vp, vp]),
    "fyzz_hip_world_step": (C.c_int, [vp, C.c_float]),
    "fyzz_hip_world_step_graph": (C.c_int, [vp, C.c_float]),
    "fyzz_hip_world_autotune": (C.c_int, [vp, C.c_float, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fyzz_hip_world_set_workgroups": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fyzz_hip_world_body_records": (C.c_int, [vp, vp, C.c_uint32, u32p]),
    "fyzz_hip_world_stats": (C.c_int, [vp, u32p, u32p, u32p, u32p]),
    "fyzz_hip_world_fetch_managers": (C.c_int, [vp, vp, vp, vp, C.c_uint32, u32p]),
    "fyzz_hip_world_friction_anchors": (C.c_int, [vp, vp, C.c_uint32, vp]),
    "fyzz_hip_world_totals": (C.c_int, [vp, u64p, C.c_int]),
    "fyzz_hip_world_set_contact_notifications": (C.c_int, [vp, C.c_int]),
    "fyzz_hip_world_contact_events": (C.c_int, [vp, vp, C.c_uint32, u32p, vp, C.c_uint32, u32p]),
    "fyzz_hip_world_broadphase_reports": (C.c_int, [vp, C.c_int, vp, C.c_uint64, u64p]),
    "fyzz_hip_world_fetch_persistent": (C.c_int, [vp, vp, C.c_uint32, u32p, vp, vp, vp, u32p]),
    "fyzz_hip_world_fetch_touch": (C.c_int, [vp, vp, C.c_uint32, vp]),
    "fyzz_hip_world_device_views": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                              C.POINTER(vp)]),
    "fyzz_hip_world_set_solver_order": (C.c_int, [vp, C.c_int]),
    "fyzz_hip_world_append_dy_inputs": (C.c_int, [vp, vp, vp, C.c_uint64, vp, vp]),
    "fyzz_hip_world_add_dynamic_sphere": (C.c_int, [vp, C.c_float, C.c_float, u32p]),
    "fyzz_hip_world_add_static_box": (C.c_int, [vp, vp, vp, u32p]),
    "fyzz_hip_world_add_static_sphere": (C.c_int, [vp, C.c_float, vp, u32p]),
    "fyzz_hip_islands_create": (C.c_int, [C.c_uint32, vp, C.POINTER(vp)]),
    "fyzz_hip_islands_destroy": (None, [vp]),
    "fyzz_hip_islands_step": (C.c_int, [vp, vp, vp]),
    "fyzz_hip_islands_last_error": (C.c_char_p, [vp]),
}


class LibraryMissing(ImportError):
    pass


class FyzzError(RuntimeError):
    """A non-zero status from the C-ABI, with the context's message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python fyzzxyzzy-sdk_amd/build.py` "
                "(there is no CPU fallback)")
        try:
            handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_LOCAL", 0))
        except OSError as e:  # pragma: no cover - reported to the caller
            raise LibraryMissing(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            f = getattr(handle, name)
            f.restype = res
            f.argtypes = args
        _lib = handle
    return _lib


def check(ctx, rc: int):
    if rc != 0:
        msg = lib().fyzz_hip_last_error(ctx)
        raise FyzzError(rc, msg.decode() if msg else "")
    return rc


class Context:
    """One device context; optionally bound to a HIP stream handle."""

    def __init__(self, device: int = 0, stream: int | None = None):
        h = vp()
        rc = lib().fyzz_hip_context_create(device, C.byref(h))
        if rc != 0:
            raise FyzzError(rc, f"fyzz_hip_context_create(device={device}) failed")
        self.h = h
        self._children = weakref.WeakSet()
        if stream is not None:
            check(self.h, lib().fyzz_hip_set_stream(self.h, vp(stream)))

    def adopt(self, child):
        """Register an object whose close() must run before this context's."""
        self._children.add(child)

    def close(self):
        if self.h:
            for c in list(self._children):
                c.close()
            lib().fyzz_hip_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name: str, *args):
        return check(self.h, getattr(lib(), name)(self.h, *args))

    def synchronize(self):
        self.call("fyzz_hip_synchronize")

    def set_cooperative(self, cooperative: bool = True):
        """Cooperative (default) or plain launches of the Dy grid-barrier
        kernels; results are identical."""
        self.call("fyzz_hip_set_cooperative", 1 if cooperative else 0)

    def profile(self, enable: bool = True):
        self.call("fyzz_hip_profile_enable", 1 if enable else 0)

    def profile_read(self, name: str):
        """(total_ms, count) of one named stage since the last read."""
        t = C.c_double()
        c = C.c_uint64()
        self.call("fyzz_hip_profile_read", name.encode(), C.byref(t), C.byref(c))
        return t.value, c.value
