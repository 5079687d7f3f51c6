"""Python side of the Dy C-ABI entries (include/fyzz_hip.h).

Record layouts mirror the header's fyzz_dy_* structs; device records are
torch uint8 tensors of shape (n, itemsize) so the C-ABI sees plain
pointers. Reference: Scene::prepare_solver_bodies .. integrate_solver_task
(src/native/scene.cpp:6940-10707).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Context, FyzzError, lib

vp = C.c_void_p

DY_BODY = np.dtype([("position", "<f4", 3), ("orientation", "<f4", 4), ("linear_velocity", "<f4", 3),
                    ("angular_velocity", "<f4", 3), ("inverse_mass", "<f4"), ("inverse_inertia", "<f4", 3),
                    ("wake_counter", "<f4"), ("sleep_linear_acc", "<f4", 3), ("sleep_angular_acc", "<f4", 3),
                    ("counted_interactions", "<u4"), ("flags", "<u4"), ("linear_damping", "<f4"),
                    ("angular_damping", "<f4"), ("max_linear_velocity", "<f4"), ("max_angular_velocity", "<f4"),
                    ("static_friction", "<f4"), ("dynamic_friction", "<f4"), ("restitution", "<f4"),
                    ("iterations", "<u4"), ("pad", "<f4", 2)])
DY_STATIC = np.dtype([("position", "<f4", 3), ("orientation", "<f4", 4), ("static_friction", "<f4"),
                      ("dynamic_friction", "<f4"), ("restitution", "<f4")])
DY_CONTACT = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("separation", "<f4")])
DY_MANAGER = np.dtype([("body0", "<i4"), ("body1", "<i4"), ("count", "<u4"), ("history", "<u4"),
                       ("contacts", DY_CONTACT, 4)])
DY_PATCH = np.dtype([("body0_normal", "<f4", 3), ("body1_normal", "<f4", 3), ("body0_anchors", "<f4", (2, 3)),
                     ("body1_anchors", "<f4", (2, 3)), ("static_friction", "<f4"), ("dynamic_friction", "<f4"),
                     ("restitution", "<f4"), ("damping", "<f4"), ("anchor_count", "<u4"), ("flags", "<u4")])
# fyzz_friction_anchors: Scene::FrictionPatchWriteback per manager
FRICTION_ANCHORS = np.dtype([("anchor_count", "<u4"), ("broken", "<u4"), ("position", "<f4", (2, 3)),
                             ("impulse", "<f8", (2, 3))])
assert FRICTION_ANCHORS.itemsize == 80
assert DY_BODY.itemsize == 144 and DY_STATIC.itemsize == 40
assert DY_MANAGER.itemsize == 128 and DY_PATCH.itemsize == 96
NO_HISTORY = 0xFFFFFFFF

FLAG_NONFINITE, FLAG_LOGIC, FLAG_OVERFLOW, FLAG_BAD_INPUT = 1, 2, 4, 8


def default_bodies(n):
    """Dynamic-box records with RigidBodyDynamics defaults (reference
    include/fyzzxyzzy/world.h:115-140): 4 position + 1 velocity
    iterations, angular damping 0.05, gravity on, awake."""
    b = np.zeros(n, DY_BODY)
    b["orientation"][:, 3] = 1.0
    b["wake_counter"] = 0.4
    b["flags"] = 4
    b["angular_damping"] = 0.05
    b["max_linear_velocity"] = 1.0e16
    b["max_angular_velocity"] = 100.0
    b["static_friction"] = 0.5
    b["dynamic_friction"] = 0.5
    b["restitution"] = 0.2
    b["iterations"] = 4 | (1 << 8)
    return b


def to_device(records, device="cuda"):
    """Structured numpy records -> (n, itemsize) uint8 device tensor."""
    import torch
    a = np.ascontiguousarray(records)
    return torch.from_numpy(a.view(np.uint8).reshape(len(a), a.dtype.itemsize).copy()).to(device)


def from_device(t, dtype):
    return np.ascontiguousarray(t.cpu().numpy()).view(dtype).reshape(-1)


def _p(t):
    return vp(t.data_ptr()) if t is not None and t.numel() else None


class DySolver:
    """Workspace for Dy steps of up to body_capacity bodies and
    manager_capacity managers on one context."""

    def __init__(self, ctx: Context, body_capacity: int, manager_capacity: int):
        self.ctx = ctx
        h = vp()
        ctx.call("fyzz_hip_dy_create", C.c_uint32(body_capacity), C.c_uint32(manager_capacity), C.byref(h))
        self.h = h
        ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None):
            lib().fyzz_hip_dy_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        rc = getattr(lib(), name)(self.h, *args)
        if rc != 0:
            msg = lib().fyzz_hip_last_error(self.ctx.h)
            raise FyzzError(rc, msg.decode() if msg else "")

    def step(self, dt, bodies, statics, managers, history=None, patches=None, impulses=None, partitions=None):
        """One step over device records (bodies updated in place). Returns
        (patches, impulses, partitions) device tensors."""
        import torch
        dev = bodies.device
        nm = managers.shape[0]
        if patches is None:
            patches = torch.zeros((max(nm, 1), DY_PATCH.itemsize), dtype=torch.uint8, device=dev)
        if impulses is None:
            impulses = torch.zeros((max(nm, 1), 4), dtype=torch.float32, device=dev)
        if partitions is None:
            partitions = torch.zeros(max(nm, 1), dtype=torch.int32, device=dev)
        nh = 0 if history is None else history.shape[0]
        self._call("fyzz_hip_dy_step", C.c_float(dt), C.c_uint32(bodies.shape[0]), _p(bodies),
                   C.c_uint32(statics.shape[0]), _p(statics), C.c_uint32(nm), _p(managers), _p(history),
                   C.c_uint32(nh), _p(patches), _p(impulses), _p(partitions))
        return patches[:nm], impulses[:nm], partitions[:nm]

    def friction_anchors(self, n_managers):
        """The last step's friction write-back (one FRICTION_ANCHORS record
        per manager), computed from its solved rows; the patches tensor that
        step wrote must still be alive. Host copy."""
        import torch
        out = torch.zeros((max(n_managers, 1), FRICTION_ANCHORS.itemsize), dtype=torch.uint8,
                          device=torch.device("cuda", 0))
        self._call("fyzz_hip_dy_friction_anchors", _p(out))
        return from_device(out, FRICTION_ANCHORS)[:n_managers]

    def set_tape(self, enabled=True):
        self._call("fyzz_hip_dy_set_tape", C.c_int(1 if enabled else 0))

    def adjoint(self, adj_v_out, adj_w_out, adj_p_out):
        """Reverse of the last (taped) step: returns (adj_v_in, adj_w_in,
        adj_dynamic_friction) device tensors."""
        import torch
        nb = adj_v_out.shape[0]
        av = torch.empty((nb, 3), dtype=torch.float32, device=adj_v_out.device)
        aw = torch.empty_like(av)
        af = torch.empty(nb, dtype=torch.float32, device=adj_v_out.device)
        c = lambda t: t.contiguous()  # noqa: E731
        a1, a2, a3 = c(adj_v_out), c(adj_w_out), c(adj_p_out)
        self._call("fyzz_hip_dy_adjoint", _p(a1), _p(a2), _p(a3), _p(av), _p(aw), _p(af))
        return av, aw, af

    def partition(self, bodies, managers):
        import torch
        nm = managers.shape[0]
        out = torch.zeros(max(nm, 1), dtype=torch.int32, device=bodies.device)
        self._call("fyzz_hip_dy_partition", C.c_uint32(bodies.shape[0]), _p(bodies), C.c_uint32(nm),
                   _p(managers), _p(out))
        return out[:nm]

    def status(self, raise_on_error=True):
        """Synchronizes; returns (flags, partitions, rounds) of the last step."""
        f, n, r = C.c_uint32(), C.c_uint32(), C.c_uint32()
        rc = lib().fyzz_hip_dy_status(self.h, C.byref(f), C.byref(n), C.byref(r))
        if rc != 0 and raise_on_error:
            msg = lib().fyzz_hip_last_error(self.ctx.h)
            raise FyzzError(rc, msg.decode() if msg else "")
        return f.value, n.value, r.value
