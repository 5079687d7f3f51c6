"""The compiled reference NativeCpu scene (oracle/_ref/libfyzzref.so, built
from the reference's own sources by oracle/Makefile) stepped live as a
checker: each step yields what a Dy replay in the reference's own solver
order needs and what it must reproduce. Test helper, not a test module;
tests that use it skip when the library was not built.

Per step s: the start state (13 floats per box: position, linear velocity,
orientation, angular velocity), wake counters, ready and sleeping flags;
the step's descriptor stream (the reference's island order) as Dy manager
records with the reference's published contacts (plane contacts with the
published +y normal, box/box pairs with the higher body first, as the
narrowphase hands them to Dy); each descriptor's partition local to its
task; the broadphase pair counts per box; and the state after the step."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
REF_LIB = ROOT / "oracle" / "_ref" / "libfyzzref.so"
P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


def available() -> bool:
    return REF_LIB.exists()


def _load():
    ref = C.CDLL(str(REF_LIB))
    ref.ref_scene_create.restype = C.c_void_p
    ref.ref_last_error.restype = C.c_char_p
    for f in ("ref_pairs", "ref_manifolds", "ref_dy_descriptors", "ref_dy_tasks"):
        getattr(ref, f).restype = C.c_size_t
    return ref


class RefRun:
    """A scenes.Scene record authored into the reference scene (floor id 0,
    boxes 1..n in order, as the device world authors it)."""

    def __init__(self, scene):
        self.ref = ref = _load()
        self.s = C.c_void_p(ref.ref_scene_create())
        self.n = n = scene.n
        fid = C.c_uint32()
        self._ok(ref.ref_add_floor(self.s, C.c_float(scene.floor_height), C.byref(fid)))
        self._ok(ref.ref_set_material(self.s, fid, C.c_float(scene.friction), C.c_float(scene.friction),
                                      C.c_float(scene.restitution)))
        first = getattr(scene, "statics_first", False)
        if first:
            self._add_statics(scene)
        self.ids = np.zeros(n, np.uint32)
        for i in range(n):
            bid = C.c_uint32()
            if scene.sphere[i]:
                self._ok(ref.ref_add_dynamic_sphere(self.s, C.c_float(float(scene.half[i, 0])),
                                                    P(np.ascontiguousarray(scene.pos[i], np.float32)),
                                                    C.c_float(scene.density), C.byref(bid)))
            else:
                self._ok(ref.ref_add_dynamic_box(self.s, P(np.ascontiguousarray(scene.half[i], np.float32)),
                                                 P(np.ascontiguousarray(scene.pos[i], np.float32)),
                                                 C.c_float(scene.density), C.byref(bid)))
            self.ids[i] = bid.value
        for i in range(n):
            b = C.c_uint32(int(self.ids[i]))
            self._ok(ref.ref_set_pose(self.s, b, P(np.ascontiguousarray(scene.pos[i], np.float32)),
                                      P(np.ascontiguousarray(scene.quat[i], np.float32))))
            self._ok(ref.ref_set_velocity(self.s, b, P(np.ascontiguousarray(scene.vel[i], np.float32))))
            if np.any(scene.ang[i]):
                self._ok(ref.ref_set_angular_velocity(self.s, b, P(np.ascontiguousarray(scene.ang[i], np.float32))))
            self._ok(ref.ref_set_material(self.s, b, C.c_float(scene.friction), C.c_float(scene.friction),
                                          C.c_float(scene.restitution)))
        base = 1 + (len(scene.statics) if first else 0)
        assert list(self.ids) == list(range(base, base + n))
        if not first:
            self._add_statics(scene)

    def _add_statics(self, scene):
        ref = self.ref
        for h, pos, quat, is_sphere in scene.statics:
            sid = C.c_uint32()
            p = P(np.ascontiguousarray(pos, np.float32))
            if is_sphere:
                self._ok(ref.ref_add_static_sphere(self.s, C.c_float(float(h)), p, C.byref(sid)))
            else:
                self._ok(ref.ref_add_static_box(self.s, P(np.ascontiguousarray(h, np.float32)), p, C.byref(sid)))
            self._ok(ref.ref_set_pose(self.s, sid, p, P(np.ascontiguousarray(quat, np.float32))))
            self._ok(ref.ref_set_material(self.s, sid, C.c_float(scene.friction), C.c_float(scene.friction),
                                          C.c_float(scene.restitution)))

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(self.ref.ref_last_error().decode())

    def close(self):
        if self.s:
            self.ref.ref_scene_destroy(self.s)
            self.s = None

    def set_contact_notifications(self, enabled=True):
        self._ok(self.ref.ref_set_contact_notifications(self.s, 1 if enabled else 0))

    def contact_events(self):
        """The last step's published events: ev[k, 6] (body0, body1, kind,
        count, shape0, shape1), offsets, f64 impulses; points[m, 7]
        (position, normal, separation), f64 point impulses."""
        ref = self.ref
        ref.ref_contact_events.restype = C.c_size_t
        npts = C.c_size_t()
        ne = ref.ref_contact_events(self.s, None, None, None, 0, None, None, 0, C.byref(npts))
        ev = np.zeros((max(ne, 1), 6), np.uint32)
        off = np.zeros(max(ne, 1), np.uint64)
        imp = np.zeros(max(ne, 1), np.float64)
        pt = np.zeros((max(npts.value, 1), 7), np.float32)
        pimp = np.zeros(max(npts.value, 1), np.float64)
        ref.ref_contact_events(self.s, P(ev), P(off), P(imp), ne, P(pt), P(pimp), npts.value, C.byref(npts))
        return ev[:ne], off[:ne], imp[:ne], pt[:npts.value], pimp[:npts.value]

    def state(self):
        st = np.zeros(14, np.float32)
        out = np.zeros((self.n, 13), np.float32)
        sleeping = np.zeros(self.n, bool)
        wake = np.zeros(self.n, np.float32)
        ready = np.zeros(self.n, bool)
        for i in range(self.n):
            b = C.c_uint32(int(self.ids[i]))
            self._ok(self.ref.ref_body_state(self.s, b, P(st)))
            out[i] = st[:13]
            sleeping[i] = st[13] > 0
            w, r = C.c_float(), C.c_int()
            self._ok(self.ref.ref_sleep_state(self.s, b, C.byref(w), C.byref(r)))
            wake[i], ready[i] = w.value, bool(r.value)
        return out, wake, ready, sleeping

    def step(self, dt):
        """Steps once; returns the step's Dy inputs and outputs (a dict)."""
        from fyzz_amd.dy import DY_MANAGER
        ref, n = self.ref, self.n
        state, wake, ready, sleeping = self.state()
        self._ok(ref.ref_step(self.s, C.c_float(dt)))
        m = ref.ref_manifolds(self.s, None, None, None, 0)
        mp = np.zeros(2 * max(m, 1), np.uint32)
        mc = np.zeros(max(m, 1), np.uint8)
        pts = np.zeros(max(m, 1) * 28, np.float32)
        ref.ref_manifolds(self.s, P(mp), P(mc), P(pts), m)
        mp, pts = mp.reshape(-1, 2), pts.reshape(-1, 4, 7)
        d = ref.ref_dy_descriptors(self.s, None, None, None, None, 0)
        di = np.zeros(max(d, 1), np.uint32)
        b0 = np.zeros(max(d, 1), np.uint32)
        b1 = np.zeros(max(d, 1), np.uint32)
        part = np.zeros(max(d, 1), np.uint64)
        ref.ref_dy_descriptors(self.s, P(di), P(b0), P(b1), P(part), d)
        t = ref.ref_dy_tasks(self.s, None, None, 0)
        me = np.zeros(max(t, 1), np.uint64)
        pe = np.zeros(max(t, 1), np.uint64)
        ref.ref_dy_tasks(self.s, P(me), P(pe), t)
        mg = np.zeros(d, DY_MANAGER)
        keys = []
        for k in range(d):
            a, b = (int(x) for x in mp[di[k]])
            cnt = int(mc[di[k]])
            if a == 0:  # the floor (body 0): the box first, the static second
                mg[k]["body0"], mg[k]["body1"] = b - 1, -1
            else:  # two boxes: the higher body first
                mg[k]["body0"], mg[k]["body1"] = b - 1, a - 1
            mg[k]["count"] = cnt
            mg[k]["history"] = 0xFFFFFFFF
            for c in range(cnt):
                mg[k]["contacts"][c]["position"] = pts[di[k], c, 0:3]
                mg[k]["contacts"][c]["normal"] = pts[di[k], c, 3:6]
                mg[k]["contacts"][c]["separation"] = pts[di[k], c, 6]
            keys.append((a, b))
        # partitions local to each task
        glob = part[:d].astype(np.int64)
        ends = pe[:t].astype(np.int64)
        base = np.concatenate([[0], ends[:-1]]) if t else np.zeros(0, np.int64)
        task = np.searchsorted(ends, glob, side="right")
        local = glob - base[task] if d else glob
        np_ = ref.ref_pairs(self.s, None, 0)
        pr = np.zeros(2 * max(np_, 1), np.uint32)
        ref.ref_pairs(self.s, P(pr), np_)
        pr = pr[:2 * np_].reshape(-1, 2).astype(np.int64)
        counted = np.bincount(pr[pr > 0] - 1, minlength=n).astype(np.uint32)
        box = (pr[:, 0] > 0) & (pr[:, 1] > 0)
        pairs = sorted((int(a) - 1, int(b) - 1) for a, b in pr[box])
        live = mc[:m] > 0
        raw_pairs = pr.astype(np.uint32)
        raw_touch = mp[:m][live].astype(np.uint32)
        touch_bb = {(int(a) - 1, int(b) - 1) for (a, b) in mp[:m][live] if a > 0}
        touch_pl = {int(b) - 1 for (a, b) in mp[:m][live] if a == 0}
        after = self.state()
        return {"state": state, "wake": wake, "ready": ready, "sleeping": sleeping, "managers": mg, "keys": keys,
                "partitions": local, "tasks": t, "counted": counted, "next": after, "pairs": pairs,
                "touch_bb": touch_bb, "touch_pl": touch_pl, "raw_pairs": raw_pairs, "raw_touch": raw_touch,
                "task_manager_ends": me[:t].astype(np.int64)}


def dy_bodies_from(step, prev, half, density=1000.0, restitution=0.0, acc_lin=None, acc_ang=None, friction=0.5):
    """Dy body records of a RefRun step: its start state (sleep accumulators
    carried by the caller from its own previous step) after the step's
    island pass (the oracle's, over the reference's pairs and touches of
    this step and prev, the previous step's dict or None), which the
    reference runs inside the step before Dy."""
    from fyzz_amd import scenes
    from oracle_bind import dy_bodies, island_pass
    st = step["state"]
    n = len(st)
    inv_mass, inv_inertia = scenes.box_mass_properties(half, density)
    b = dy_bodies(n)
    b["position"] = st[:, 0:3]
    b["linear_velocity"] = st[:, 3:6]
    b["orientation"] = st[:, 6:10]
    b["angular_velocity"] = st[:, 10:13]
    b["inverse_mass"] = inv_mass
    b["inverse_inertia"] = inv_inertia
    b["wake_counter"] = step["wake"]
    b["flags"] = 4 | (step["ready"].astype(np.uint32) << 1) | step["sleeping"].astype(np.uint32)
    if acc_lin is not None:
        b["sleep_linear_acc"] = acc_lin
        b["sleep_angular_acc"] = acc_ang
    b["counted_interactions"] = step["counted"]
    b["restitution"] = restitution
    b["static_friction"] = friction
    b["dynamic_friction"] = friction
    pp, pbb, ppl = (prev["pairs"], prev["touch_bb"], prev["touch_pl"]) if prev else ([], set(), set())
    island_pass(b, step["pairs"], pp, step["touch_bb"], pbb, step["touch_pl"], ppl)
    return b


def statics_of(scene):
    """The floor as the Dy static record the device world builds."""
    from fyzz_amd.dy import DY_STATIC
    st = np.zeros(1, DY_STATIC)
    st["orientation"][0, 3] = 1.0
    st["position"][0, 1] = scene.floor_height
    st["static_friction"] = scene.friction
    st["dynamic_friction"] = scene.friction
    st["restitution"] = scene.restitution
    return st


def replay(scene, steps, solve, dt=1.0 / 120.0):
    """Steps the reference scene and hands each step's Dy inputs, in the
    reference's descriptor order from its synchronized start state, to
    solve(s, step, bodies, statics, managers, history) -> (bodies_out,
    patches); friction history (by shape pair) and the sleep accumulators
    are carried from solve's results."""
    from fyzz_amd.dy import DY_PATCH
    run = RefRun(scene)
    statics = statics_of(scene)
    prev, history = None, {}
    acc_lin = np.zeros((scene.n, 3), np.float32)
    acc_ang = np.zeros((scene.n, 3), np.float32)
    try:
        for s in range(steps):
            st = run.step(dt)
            b = dy_bodies_from(st, prev, scene.half, scene.density, scene.restitution, acc_lin, acc_ang,
                               scene.friction)
            mg = st["managers"].copy()
            hist = []
            for i, k in enumerate(st["keys"]):
                if k in history:
                    mg[i]["history"] = len(hist)
                    hist.append(history[k])
            hist = np.array(hist, DY_PATCH) if hist else np.zeros(0, DY_PATCH)
            out, patches = solve(s, st, b, statics, mg, hist)
            history = {k: patches[i] for i, k in enumerate(st["keys"])}
            acc_lin = out["sleep_linear_acc"].copy()
            acc_ang = out["sleep_angular_acc"].copy()
            prev = st
    finally:
        run.close()


def hub_scene(side=8, seed=None):
    """A slab (half extents 4.6 x 0.25 x 4.6) resting on the floor, a
    side x side layer of boxes (half 0.45, seeded yaw and small velocities)
    on it, and a (side-1)^2 layer offset by half a cell on top, each upper
    box across four lower ones: the slab carries side^2 box/box contacts,
    more partitions than a 32-bit colour mask holds (the overflow waves)."""
    from fyzz_amd import scenes
    seed = scenes.SEED_BASE + 31 if seed is None else seed
    sp = 1.15
    g = np.arange(side) - (side - 1) / 2
    u = np.arange(side - 1) - (side - 2) / 2
    lower = np.stack(np.meshgrid(g * sp, [0.5 + 0.45 + 0.03], g * sp, indexing="ij"), -1).reshape(-1, 3)
    upper = np.stack(np.meshgrid(u * sp, [0.5 + 0.9 + 0.45 + 0.06], u * sp, indexing="ij"), -1).reshape(-1, 3)
    n = len(lower) + len(upper)
    half = np.concatenate([[[4.6, 0.25, 4.6]], np.full((n, 3), 0.45)])
    pos = np.concatenate([[[0.0, 0.25, 0.0]], lower, upper])
    yaw = np.concatenate([[0.0], (scenes.uniform(seed, 0, n) - 0.5) * 0.3])
    quat = scenes.normalize_quat32(scenes.yaw_quat(yaw))
    vel = np.concatenate([[[0.0, 0.0, 0.0]], np.stack([scenes.normal(seed, 1 + a, n) * 0.2 for a in range(3)], 1)])
    return scenes.Scene(half, pos, quat, vel)


def wake_scene(seed=None):
    """A 5x5 two-layer stack of unit boxes on the floor that settles and
    sleeps, and 12 smaller boxes dropped onto it from 6-40 m."""
    from fyzz_amd import scenes
    seed = scenes.SEED_BASE + 60 if seed is None else seed
    side = 5
    idx = np.arange(side * side * 2)
    i, k, j = idx % side, (idx // side) % side, idx // (side * side)
    pos = np.stack([(i - 2) * 1.05, 0.5 + j * 1.0, (k - 2) * 1.05], 1)
    half = np.full((len(idx), 3), 0.5)
    m = 12
    top = np.stack([(scenes.uniform(seed, 1, m) - 0.5) * 4, 6 + 34 * scenes.uniform(seed, 2, m),
                    (scenes.uniform(seed, 3, m) - 0.5) * 4], 1)
    pos = np.concatenate([pos, top])
    half = np.concatenate([half, np.stack([0.3 + 0.2 * scenes.uniform(seed, 4 + a, m) for a in range(3)], 1)])
    yaw = np.concatenate([np.zeros(len(idx)), scenes.uniform(seed, 8, m) * 2 * np.pi])
    return scenes.Scene(half, pos, scenes.normalize_quat32(scenes.yaw_quat(yaw)))


def mixed_scene(side=6, seed=None):
    """side^3 bodies on a grid 1.1 m apart above the floor, alternately
    spheres (radius U[0.3, 0.5]) and boxes (half extents U[0.3, 0.5], seeded
    yaw), with small seeded velocities: they land and pile up, touching as
    sphere/sphere, sphere/box, box/box, sphere/plane and box/plane pairs."""
    from fyzz_amd import scenes
    seed = scenes.SEED_BASE + 61 if seed is None else seed
    n = side ** 3
    idx = np.arange(n)
    i, k, j = idx % side, (idx // side) % side, idx // (side * side)
    pos = np.stack([(i - (side - 1) / 2) * 1.1, 0.6 + j * 1.1, (k - (side - 1) / 2) * 1.1], 1)
    sphere = (i + j + k) % 2 == 0
    half = np.stack([0.3 + 0.2 * scenes.uniform(seed, a, n) for a in range(3)], 1)
    half[sphere] = half[sphere][:, :1]
    yaw = np.where(sphere, 0.0, scenes.uniform(seed, 3, n) * 2 * np.pi)
    vel = np.stack([scenes.normal(seed, 4 + a, n) * 0.3 for a in range(3)], 1)
    return scenes.Scene(half, pos, scenes.normalize_quat32(scenes.yaw_quat(yaw)), vel, sphere=sphere)


def sphere_cases():
    """Isolated groups 12 m apart, each one sphere contact case at rest,
    for comparing the first step's manifolds: spheres embedded in a box
    (center nearest the +y face, the -x face, the exact center of a box
    and of a cube, where exit faces tie), in a rotated box, touching a rotated box's edge
    from outside, coincident sphere centers, spheres 0.03 m apart, a sphere
    on and one through the floor, each box case with the box created both
    before and after its sphere (both Gu orders)."""
    from fyzz_amd import scenes
    half, pos, quat, sphere = [], [], [], []
    ident = [0.0, 0.0, 0.0, 1.0]
    tilt = scenes.normalize_quat32(np.array([[0.2, 0.3, 0.1, 0.93]]))[0].tolist()
    g = [0]

    def body(h, p, q=ident, s=False):
        half.append([h, h, h] if np.isscalar(h) else list(h))
        pos.append([p[0] + 12.0 * g[0], p[1], p[2]])
        quat.append(q)
        sphere.append(s)

    def group(*bodies):
        for b in bodies:
            body(*b)
        g[0] += 1

    box = (0.5, 0.6, 0.7)
    for box_first in (True, False):
        for center, q in (((0.1, 3.5, 0.05), ident), ((-0.4, 3.1, 0.1), ident), ((0.0, 3.0, 0.0), ident),
                          ((0.15, 2.9, -0.2), tilt), ((0.62, 3.72, 0.0), tilt)):
            b = (box, (0.0, 3.0, 0.0), q, False)
            s = (0.3, center, ident, True)
            group(*((b, s) if box_first else (s, b)))
    for box_first in (True, False):  # a cube's center: all exit faces tie
        b = ((0.5, 0.5, 0.5), (0.0, 3.0, 0.0), ident, False)
        s = (0.3, (0.0, 3.0, 0.0), ident, True)
        group(*((b, s) if box_first else (s, b)))
    group((0.4, (0.0, 5.0, 0.0), ident, True), (0.25, (0.0, 5.0, 0.0), ident, True))
    group((0.4, (0.0, 5.0, 0.0), ident, True), (0.25, (0.0, 5.68, 0.0), ident, True))
    group((0.4, (0.0, 0.41, 0.0), ident, True))
    group((0.4, (0.0, 0.3, 0.0), ident, True))
    return scenes.Scene(np.array(half), np.array(pos), np.array(quat, np.float32), sphere=np.array(sphere))


def statics_scene(seed=None, first=False):
    """Dynamic spheres and boxes dropped onto static bodies: a tilted
    static slab on four legs' worth of height (half extents 3 x 0.25 x 3 at
    y = 2, turned 6 degrees about z) and a static sphere of radius 1.5
    resting in the floor's plane beside it; 48 bodies above the slab and 16
    above the sphere, alternately spheres and boxes. first: the statics are
    created before the dynamic bodies (lower ids: the other Gu order)."""
    from fyzz_amd import scenes
    seed = scenes.SEED_BASE + 62 if seed is None else seed
    a = np.arange(48)
    i, k, j = a % 4, (a // 4) % 4, a // 16
    pos1 = np.stack([(i - 1.5) * 1.2, 3.2 + j * 1.1, (k - 1.5) * 1.2], 1)
    b = np.arange(16)
    pos2 = np.stack([7.0 + (b % 4 - 1.5) * 0.6, 3.6 + (b // 4) * 1.1, (b % 2 - 0.5) * 0.5], 1)
    pos = np.concatenate([pos1, pos2])
    n = len(pos)
    sphere = np.arange(n) % 2 == 0
    half = np.stack([0.25 + 0.2 * scenes.uniform(seed, c, n) for c in range(3)], 1)
    half[sphere] = half[sphere][:, :1]
    yaw = np.where(sphere, 0.0, scenes.uniform(seed, 3, n) * 2 * np.pi)
    sc = scenes.Scene(half, pos, scenes.normalize_quat32(scenes.yaw_quat(yaw)), sphere=sphere)
    t = 6.0 * np.pi / 180.0
    slab_q = scenes.normalize_quat32(np.array([[0.0, 0.0, np.sin(t / 2), np.cos(t / 2)]]))[0]
    sc.statics = [((3.0, 0.25, 3.0), (0.0, 2.0, 0.0), slab_q, False),
                  (1.5, (7.0, 1.5, 0.0), np.float32([0, 0, 0, 1]), True)]
    sc.statics_first = first  # created before the dynamic bodies (lower ids) instead of after
    return sc
