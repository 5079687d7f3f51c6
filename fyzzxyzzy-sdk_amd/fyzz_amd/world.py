"""World: the whole step pipeline on the device (C-ABI fyzz_hip_world_*).

Mirrors the authoring and stepping surface of the reference's fyzz::World
(include/fyzzxyzzy/world.h:400-470) for a floor plane and dynamic boxes:
add_static_floor, add_dynamic_box, set_body_pose, set_body_velocity,
set_body_angular_velocity, set_body_material, set_body_dynamics, step,
body_state (the setters work before and after the first step).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Context, FyzzError, lib
from .dy import DY_BODY, DY_MANAGER, FRICTION_ANCHORS

# fyzz_contact_event / fyzz_contact_event_point (include/fyzz_hip.h), the
# layouts of fyzz::ContactEvent / ContactEventPoint
CONTACT_EVENT = np.dtype([("body0", "<u4"), ("body1", "<u4"), ("kind", "u1"), ("contact_count", "u1"),
                          ("pad", "u1", 6), ("point_offset", "<u8"), ("normal_impulse", "<f8"), ("shape0", "<u4"),
                          ("shape1", "<u4")])
CONTACT_EVENT_POINT = np.dtype([("position", "<f4", 3), ("normal", "<f4", 3), ("separation", "<f4"),
                                ("pad", "<f4"), ("normal_impulse", "<f8")])
assert CONTACT_EVENT.itemsize == 40 and CONTACT_EVENT_POINT.itemsize == 40
# fyzz_body_dynamics, the layout of fyzz::RigidBodyDynamics
BODY_DYNAMICS = np.dtype([("linear_damping", "<f4"), ("angular_damping", "<f4"), ("max_linear_velocity", "<f4"),
                          ("max_angular_velocity", "<f4"), ("max_depenetration_velocity", "<f4"),
                          ("max_contact_impulse", "<f4"), ("contact_slop_coefficient", "<f4"),
                          ("gravity_enabled", "u1"), ("pad0", "u1", 3), ("lock_flags", "<u4"),
                          ("gyroscopic_forces_enabled", "u1"), ("retain_accelerations", "u1"),
                          ("position_iterations", "u1"), ("velocity_iterations", "u1")])
assert BODY_DYNAMICS.itemsize == 40
BODY_DYNAMICS_DEFAULTS = {"linear_damping": 0.0, "angular_damping": 0.05, "max_linear_velocity": 1.0e16,
                          "max_angular_velocity": 100.0, "max_depenetration_velocity": 1.0e32,
                          "max_contact_impulse": np.finfo(np.float32).max, "contact_slop_coefficient": 0.0,
                          "gravity_enabled": 1, "lock_flags": 0, "gyroscopic_forces_enabled": 0,
                          "retain_accelerations": 0, "position_iterations": 4, "velocity_iterations": 1}

vp = C.c_void_p


def _f32(a, n):
    a = np.ascontiguousarray(a, np.float32).reshape(n)
    return a, a.ctypes.data_as(vp)


class World:
    def __init__(self, ctx: Context, max_bodies: int):
        self.ctx = ctx
        h = vp()
        ctx.call("fyzz_hip_world_create", C.c_uint32(max_bodies), C.byref(h))
        self.h = h
        self.n_dynamic = 0
        ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None):
            lib().fyzz_hip_world_destroy(self.h)
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

    def add_static_floor(self, height: float = 0.0) -> int:
        out = C.c_uint32()
        self._call("fyzz_hip_world_add_static_floor", C.c_float(height), C.byref(out))
        return out.value

    def add_dynamic_box(self, half_extents, density: float = 1000.0) -> int:
        h, hp = _f32(half_extents, 3)
        out = C.c_uint32()
        self._call("fyzz_hip_world_add_dynamic_box", hp, C.c_float(density), C.byref(out))
        self.n_dynamic += 1
        return out.value

    def add_static_box(self, half_extents, position) -> int:
        h, hp = _f32(half_extents, 3)
        p, pp = _f32(position, 3)
        out = C.c_uint32()
        self._call("fyzz_hip_world_add_static_box", hp, pp, C.byref(out))
        return out.value

    def add_static_sphere(self, radius: float, position) -> int:
        p, pp = _f32(position, 3)
        out = C.c_uint32()
        self._call("fyzz_hip_world_add_static_sphere", C.c_float(radius), pp, C.byref(out))
        return out.value

    def add_dynamic_sphere(self, radius: float, density: float = 1000.0) -> int:
        out = C.c_uint32()
        self._call("fyzz_hip_world_add_dynamic_sphere", C.c_float(radius), C.c_float(density), C.byref(out))
        self.n_dynamic += 1
        return out.value

    def add_dynamic_boxes(self, half_extents, positions, orientations, linear=None, angular=None,
                          density=1000.0, friction=0.5, restitution=0.0) -> int:
        """Bulk add_dynamic_box + set_body_pose/velocity/material; returns
        the first id (ids are consecutive)."""
        n = len(half_extents)
        h = np.ascontiguousarray(half_extents, np.float32).reshape(n, 3)
        p = np.ascontiguousarray(positions, np.float32).reshape(n, 3)
        q = np.ascontiguousarray(orientations, np.float32).reshape(n, 4)
        lv = None if linear is None else np.ascontiguousarray(linear, np.float32).reshape(n, 3)
        av = None if angular is None else np.ascontiguousarray(angular, np.float32).reshape(n, 3)
        ptr = lambda a: a.ctypes.data_as(vp) if a is not None else None  # noqa: E731
        out = C.c_uint32()
        self._call("fyzz_hip_world_add_dynamic_boxes", C.c_uint32(n), ptr(h), C.c_float(density), ptr(p), ptr(q),
                   ptr(lv), ptr(av), C.c_float(friction), C.c_float(friction), C.c_float(restitution),
                   C.byref(out))
        self.n_dynamic += n
        return out.value

    def set_materials(self, first_id, static_friction, dynamic_friction, restitution):
        sf = np.ascontiguousarray(static_friction, np.float32)
        df = np.ascontiguousarray(dynamic_friction, np.float32)
        rs = np.ascontiguousarray(restitution, np.float32)
        self._call("fyzz_hip_world_set_materials", C.c_uint32(first_id), C.c_uint32(len(sf)), sf.ctypes.data_as(vp),
                   df.ctypes.data_as(vp), rs.ctypes.data_as(vp))

    def record_dy_inputs(self):
        """The last step's (managers, history) as device uint8 tensors of
        DY_MANAGER / DY_PATCH records (history indexed by manager)."""
        import torch
        from .dy import DY_PATCH
        cnt = C.c_uint32()
        self._call("fyzz_hip_world_record_dy_inputs", None, None, C.c_uint32(0), C.byref(cnt))
        n = cnt.value
        mg = torch.empty((max(n, 1), DY_MANAGER.itemsize), dtype=torch.uint8, device="cuda")
        hs = torch.empty((max(n, 1), DY_PATCH.itemsize), dtype=torch.uint8, device="cuda")
        self._call("fyzz_hip_world_record_dy_inputs", vp(mg.data_ptr()), vp(hs.data_ptr()), C.c_uint32(max(n, 1)),
                   C.byref(cnt))
        return mg[:n], hs[:n]

    def append_dy_inputs(self, managers, history, capacity, cursor, count):
        """Appends the last step's Dy inputs to device arenas (uint8 tensors of
        DY_MANAGER / DY_PATCH records) at cursor (int64 device tensor),
        writing the step's count (int32 device tensor); no host round trip."""
        self._call("fyzz_hip_world_append_dy_inputs", vp(managers.data_ptr()), vp(history.data_ptr()),
                   C.c_uint64(capacity), vp(cursor.data_ptr()), vp(count.data_ptr()))

    def copy_bodies(self):
        """Device copy of the body records, (n, 144) uint8 tensor."""
        import torch
        out = torch.empty((max(self.n_dynamic, 1), DY_BODY.itemsize), dtype=torch.uint8, device="cuda")
        self._call("fyzz_hip_world_copy_bodies", vp(out.data_ptr()))
        return out[:self.n_dynamic]

    def set_body_pose(self, body, position, orientation):
        p, pp = _f32(position, 3)
        q, qp = _f32(orientation, 4)
        self._call("fyzz_hip_world_set_body_pose", C.c_uint32(body), pp, qp)

    def set_body_velocity(self, body, linear):
        v, vpp = _f32(linear, 3)
        self._call("fyzz_hip_world_set_body_velocity", C.c_uint32(body), vpp, None)

    def set_body_angular_velocity(self, body, angular):
        w, wp = _f32(angular, 3)
        self._call("fyzz_hip_world_set_body_velocity", C.c_uint32(body), None, wp)

    def set_body_material(self, body, static_friction, dynamic_friction, restitution):
        self._call("fyzz_hip_world_set_body_material", C.c_uint32(body), C.c_float(static_friction),
                   C.c_float(dynamic_friction), C.c_float(restitution))

    def set_body_dynamics(self, body, **fields):
        """World::set_body_dynamics: RigidBodyDynamics defaults
        (world.h:115-140) overridden by keyword."""
        d = np.zeros(1, BODY_DYNAMICS)
        for k, v in BODY_DYNAMICS_DEFAULTS.items():
            d[k] = v
        for k, v in fields.items():
            if k not in BODY_DYNAMICS_DEFAULTS:
                raise TypeError(f"unknown dynamics field {k}")
            d[k] = v
        self._call("fyzz_hip_world_set_body_dynamics", C.c_uint32(body), d.ctypes.data_as(vp))

    def set_scenes(self, box_counts):
        """Independent scenes: scene k owns the next box_counts[k] boxes."""
        c = np.ascontiguousarray(box_counts, np.uint32)
        self._call("fyzz_hip_world_set_scenes", C.c_uint32(len(c)), c.ctypes.data_as(vp))

    def step(self, dt: float):
        self._call("fyzz_hip_world_step", C.c_float(dt))

    def step_graph(self, dt: float):
        """A step through the captured graphs (worlds with scenes)."""
        self._call("fyzz_hip_world_step_graph", C.c_float(dt))

    def autotune(self, dt: float, trial_steps: int = 3):
        """Times real steps per launch shape and keeps the fastest; returns
        (partition workgroups, PGS workgroups)."""
        pb, sb = C.c_int(), C.c_int()
        self._call("fyzz_hip_world_autotune", C.c_float(dt), C.c_uint32(trial_steps), C.byref(pb), C.byref(sb))
        return pb.value, sb.value

    def set_workgroups(self, partition_blocks: int = 0, pgs_blocks: int = 0):
        """Sets the partition and PGS workgroup counts (0 keeps one), clamped
        to what stays co-resident; returns the counts in effect."""
        pb, sb = C.c_int(), C.c_int()
        self._call("fyzz_hip_world_set_workgroups", C.c_int(partition_blocks), C.c_int(pgs_blocks), C.byref(pb),
                   C.byref(sb))
        return pb.value, sb.value

    def set_solver_order(self, native: bool):
        """Before the first step: False (default) the device order, True the
        native scene's island order from the host island authority (eager
        steps only, two synchronizations per step)."""
        self._call("fyzz_hip_world_set_solver_order", C.c_int(1 if native else 0))

    def body_records(self):
        out = np.zeros(max(self.n_dynamic, 1), DY_BODY)
        cnt = C.c_uint32()
        self._call("fyzz_hip_world_body_records", out.ctypes.data_as(vp), C.c_uint32(len(out)), C.byref(cnt))
        return out[:cnt.value]

    def body_states(self):
        """(n, 13): position, linear velocity, orientation, angular velocity."""
        b = self.body_records()
        return np.concatenate([b["position"], b["linear_velocity"], b["orientation"], b["angular_velocity"]], 1)

    def stats(self):
        a, b, c, d = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._call("fyzz_hip_world_stats", C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"pairs": a.value, "managers": b.value, "partitions": c.value, "flags": d.value}

    def set_contact_notifications(self, enabled=True):
        self._call("fyzz_hip_world_set_contact_notifications", C.c_int(1 if enabled else 0))

    def contact_events(self):
        """(events, points) of the last step, host copies."""
        ne, npt = C.c_uint32(), C.c_uint32()
        self._call("fyzz_hip_world_contact_events", None, C.c_uint32(0), C.byref(ne), None, C.c_uint32(0),
                   C.byref(npt))
        ev = np.zeros(max(ne.value, 1), CONTACT_EVENT)
        pt = np.zeros(max(npt.value, 1), CONTACT_EVENT_POINT)
        self._call("fyzz_hip_world_contact_events", ev.ctypes.data_as(vp), C.c_uint32(ne.value), C.byref(ne),
                   pt.ctypes.data_as(vp), C.c_uint32(npt.value), C.byref(npt))
        return ev[:ne.value], pt[:npt.value]

    def totals(self, reset=False):
        """Running totals: steps, pairs, managers, contact rows, friction
        rows, partitions, solver iterations, partition rounds, broadphase
        found + lost pairs (summed over steps)."""
        out = (C.c_uint64 * 9)()
        self._call("fyzz_hip_world_totals", out, C.c_int(1 if reset else 0))
        keys = ("steps", "pairs", "managers", "contact_rows", "friction_rows", "partitions", "iterations", "rounds",
                "found_lost")
        return {k: int(out[i]) for i, k in enumerate(keys)}

    def broadphase_reports(self, which=0):
        """Sorted (lo, hi) slot pairs of the last step: 0 live pairs, 1
        found, 2 lost (host copy)."""
        written = C.c_uint64()
        self._call("fyzz_hip_world_broadphase_reports", C.c_int(which), None, C.c_uint64(0), C.byref(written))
        out = np.zeros((max(written.value, 1), 2), np.uint32)
        self._call("fyzz_hip_world_broadphase_reports", C.c_int(which), out.ctypes.data_as(vp),
                   C.c_uint64(len(out)), C.byref(written))
        return out[:written.value]

    def persistent_state(self):
        """Host copies of what the next step reads besides the body records:
        {"caches_bb": [P, 56] f32 aligned with the pair list, "caches_pl":
        [n, 56] f32, "history": DY_PATCH[P + n], "stamps": u32[P + n],
        "step": steps taken, "touch_bb": u32[P], "touch_pl": u32[n] (the
        last step's touching flags)}."""
        from .dy import DY_PATCH
        np_ = C.c_uint32()
        step = C.c_uint32()
        self._call("fyzz_hip_world_fetch_persistent", None, C.c_uint32(0), C.byref(np_), None, None, None,
                   C.byref(step))
        p, n = np_.value, self.n_dynamic
        cbb = np.zeros((max(p, 1), 56), np.float32)
        cpl = np.zeros((max(n, 1), 56), np.float32)
        hist = np.zeros(p + n, DY_PATCH)
        stamps = np.zeros(p + n, np.uint32)
        self._call("fyzz_hip_world_fetch_persistent", cbb.ctypes.data_as(vp), C.c_uint32(max(p, 1)), C.byref(np_),
                   cpl.ctypes.data_as(vp), hist.ctypes.data_as(vp), stamps.ctypes.data_as(vp), C.byref(step))
        assert np_.value == p
        tbb = np.zeros(max(p, 1), np.uint32)
        tpl = np.zeros(max(n, 1), np.uint32)
        self._call("fyzz_hip_world_fetch_touch", tbb.ctypes.data_as(vp), C.c_uint32(len(tbb)),
                   tpl.ctypes.data_as(vp))
        return {"caches_bb": cbb[:p], "caches_pl": cpl[:n], "history": hist, "stamps": stamps, "step": step.value,
                "touch_bb": tbb[:p], "touch_pl": tpl[:n]}

    def friction_anchors(self):
        """The last step's friction write-back (Scene::solved_friction_patches),
        one FRICTION_ANCHORS record per manager in manager order."""
        cnt = C.c_uint32()
        self._call("fyzz_hip_world_friction_anchors", None, C.c_uint32(0), C.byref(cnt))
        out = np.zeros(max(cnt.value, 1), FRICTION_ANCHORS)
        self._call("fyzz_hip_world_friction_anchors", out.ctypes.data_as(vp), C.c_uint32(len(out)), C.byref(cnt))
        return out[:cnt.value]

    def managers(self):
        """(managers, impulses[n, 4], partitions[n]) of the last step, host copies."""
        cnt = C.c_uint32()
        self._call("fyzz_hip_world_fetch_managers", None, None, None, C.c_uint32(0), C.byref(cnt))
        k = cnt.value
        mg = np.zeros(max(k, 1), DY_MANAGER)
        im = np.zeros((max(k, 1), 4), np.float32)
        pa = np.zeros(max(k, 1), np.uint32)
        self._call("fyzz_hip_world_fetch_managers", mg.ctypes.data_as(vp), im.ctypes.data_as(vp),
                   pa.ctypes.data_as(vp), C.c_uint32(k), C.byref(cnt))
        return mg[:k], im[:k], pa[:k]
