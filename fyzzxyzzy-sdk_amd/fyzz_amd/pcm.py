"""Python side of the PCM C-ABI entries (device buffers are torch tensors).

The persistent manifold cache of a pair lives in a caller-owned device
array of fyzz_hip_pcm_cache_bytes() records; a zero record is an empty
cache (Scene::PcmManifoldCache default, reference src/native/scene.h:1448).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Context, lib

vp = C.c_void_p
CACHE_FLOATS = 56


def _p(t):
    return vp(t.data_ptr()) if t is not None else None


def cache_bytes() -> int:
    return int(lib().fyzz_hip_pcm_cache_bytes())


def new_caches(n: int, device):
    import torch
    assert cache_bytes() == 4 * CACHE_FLOATS
    return torch.zeros((n, CACHE_FLOATS), dtype=torch.float32, device=device)


def box_box(ctx: Context, first, second, pos, quat, half, cache, contact_distance=0.04, cache_index=None):
    """Persistent box/box manifolds; returns (contacts[n,4,7], counts[n], flags[n])."""
    import torch
    n = first.numel()
    dev = pos.device
    contacts = torch.zeros((n, 4, 7), dtype=torch.float32, device=dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    flags = torch.zeros(n, dtype=torch.int32, device=dev)
    scratch = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    ctx.call("fyzz_hip_pcm_box_box", C.c_uint32(n), _p(first), _p(second), _p(cache_index), _p(pos), _p(quat),
             _p(half), C.c_float(contact_distance), _p(cache), _p(contacts), _p(counts), _p(flags), _p(scratch))
    return contacts, counts, flags


def plane_box(ctx: Context, box_slot, pos, quat, half, cache, plane_position=(0.0, 0.0, 0.0),
              contact_distance=0.04, cache_index=None):
    import torch
    n = box_slot.numel()
    dev = pos.device
    contacts = torch.zeros((n, 4, 7), dtype=torch.float32, device=dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    flags = torch.zeros(n, dtype=torch.int32, device=dev)
    pp = np.ascontiguousarray(plane_position, np.float32)
    ctx.call("fyzz_hip_pcm_plane_box", C.c_uint32(n), _p(box_slot), _p(cache_index), _p(pos), _p(quat), _p(half),
             pp.ctypes.data_as(vp), C.c_float(contact_distance), _p(cache), _p(contacts), _p(counts), _p(flags))
    return contacts, counts, flags


def candidates(ctx: Context, a, b, contact_distance=0.04):
    import torch
    n = a.shape[0]
    dev = a.device
    out = torch.zeros((n, 16, 10), dtype=torch.float32, device=dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    ov = torch.zeros(n, dtype=torch.int32, device=dev)
    sel = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    ctx.call("fyzz_hip_pcm_candidates", C.c_uint32(n), _p(a), _p(b), C.c_float(contact_distance), _p(out),
             _p(counts), _p(ov), _p(sel))
    return out, counts, ov, sel


def gjk_epa(ctx: Context, a, b, contact_distance=0.04):
    import torch
    n = a.shape[0]
    dev = a.device
    out = torch.zeros((n, 10), dtype=torch.float32, device=dev)
    ok = torch.zeros(n, dtype=torch.int32, device=dev)
    ctx.call("fyzz_hip_pcm_gjk_epa", C.c_uint32(n), _p(a), _p(b), C.c_float(contact_distance), _p(out), _p(ok))
    return out, ok
