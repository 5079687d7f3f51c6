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


class OraclePcm:
    """PCM entry points of the oracle (see oracle/orc_pcm.cpp)."""

    CACHE_FLOATS = 56

    def __init__(self, lib):
        self.lib = lib
        self.lib.orc_pcm_cache_floats.restype = C.c_uint32
        assert self.lib.orc_pcm_cache_floats() == self.CACHE_FLOATS

    def candidates(self, a, b, cd):
        out = np.zeros(64 * 10, np.float32)
        cnt = C.c_uint32()
        ok = C.c_int()
        rc = self.lib.orc_pcm_box_box_candidates(P(np.ascontiguousarray(a, np.float32)),
                                                 P(np.ascontiguousarray(b, np.float32)), C.c_float(cd), P(out),
                                                 C.c_uint32(64), C.byref(cnt), C.byref(ok))
        assert rc == 0
        return bool(ok.value), out.reshape(64, 10)[: cnt.value]

    def batch_selection(self, cands, tol=1.0):
        sel = np.zeros(4, np.uint8)
        assert self.lib.orc_pcm_batch_selection(P(np.ascontiguousarray(cands, np.float32)), C.c_uint32(len(cands)),
                                                C.c_float(tol), P(sel)) == 0
        return sel

    def gjk_epa(self, a, b, cd):
        out = np.zeros(10, np.float32)
        ok = C.c_int()
        self.lib.orc_pcm_gjk_epa(P(np.ascontiguousarray(a, np.float32)), P(np.ascontiguousarray(b, np.float32)),
                                 C.c_float(cd), P(out), C.byref(ok))
        return bool(ok.value), out

    def box_box(self, first, second, cd, cache):
        pts = np.zeros(4 * 7, np.float32)
        fl = C.c_uint32()
        n = self.lib.orc_pcm_box_box(P(np.ascontiguousarray(first, np.float32)),
                                     P(np.ascontiguousarray(second, np.float32)), C.c_float(cd), P(cache), P(pts),
                                     C.byref(fl))
        if n < 0:
            raise OverflowError("PCM box/box manifold invalid")
        return n, pts.reshape(4, 7)[:n], fl.value

    def plane_box(self, plane_pos, box, cd, cache):
        pts = np.zeros(4 * 7, np.float32)
        fl = C.c_uint32()
        n = self.lib.orc_pcm_plane_box(P(np.ascontiguousarray(plane_pos, np.float32)),
                                       P(np.ascontiguousarray(box, np.float32)), C.c_float(cd), P(cache), P(pts),
                                       C.byref(fl))
        if n < 0:
            raise OverflowError("PCM plane/box manifold invalid")
        return n, pts.reshape(4, 7)[:n], fl.value

    def new_cache(self):
        return np.zeros(self.CACHE_FLOATS, np.float32)


def replay_pcm_scene(orc, g, on_step=None):
    """Replays the narrowphase of a recorded reference scene with the oracle.
    Pairs come from the oracle SAP on each step's start poses (contact
    offset 0.02 per shape, the floor as a plane pair); caches persist by
    pair. Calls on_step(step, manifolds) where manifolds maps a shape pair
    to (count, points[count, 7]) with published normals (plane pairs flip
    the Gu plane-first normal, as the scene does before Dy)."""
    pcm = OraclePcm(orc.lib)
    half, ids = g["half"], g["ids"]
    floor = int(g["floor_id"][0])
    n = len(ids)
    caches = {}
    for s in range(int(g["steps"][0])):
        pose = g["poses"][s]
        pos, quat = pose[:, :3].copy(), pose[:, 3:].copy()
        b = orc.box_bounds(pos, quat, half, np.full(n, 0.02, np.float32))
        pairs = orc.sap_pairs(b, shape_id=ids, plane_ids=[floor])
        slot = {int(v): i for i, v in enumerate(ids)}
        new_caches = {}
        manifolds = {}
        for pa, pb in pairs:
            key = (int(pa), int(pb))
            cache = caches.get(key)
            if cache is None:
                cache = pcm.new_cache()
            if pa == floor or pb == floor:
                bx = slot[int(pb if pa == floor else pa)]
                rec = np.concatenate([half[bx], pos[bx], quat[bx]])
                cnt, pts, _ = pcm.plane_box(np.array([0.0, 0.0, 0.0], np.float32), rec, 0.04, cache)
                pts = pts.copy()
                pts[:, 3:6] *= -1.0
            else:
                # two dynamics: the Gu pair puts the higher body id first
                first, second = slot[int(pb)], slot[int(pa)]
                r1 = np.concatenate([half[first], pos[first], quat[first]])
                r2 = np.concatenate([half[second], pos[second], quat[second]])
                cnt, pts, _ = pcm.box_box(r1, r2, 0.04, cache)
            new_caches[key] = cache
            if cnt:
                manifolds[key] = (cnt, pts)
        caches = new_caches
        if on_step:
            on_step(s, pairs, manifolds)


from fyzz_amd.dy import DY_BODY, DY_CONTACT, DY_MANAGER, DY_PATCH, DY_STATIC  # noqa: E402,F401


def dy_bodies(n):
    """Default dynamic-box records (RigidBodyDynamics defaults, world.h:115-140)."""
    b = np.zeros(n, DY_BODY)
    b["orientation"][:, 3] = 1.0
    b["wake_counter"] = 0.4
    b["flags"] = 4  # gravity enabled
    b["angular_damping"] = 0.05
    b["max_linear_velocity"] = 1.0e16
    b["max_angular_velocity"] = 100.0
    b["static_friction"] = 0.5
    b["dynamic_friction"] = 0.5
    b["restitution"] = 0.2
    b["iterations"] = 4 | (1 << 8)
    return b


class OracleDy:
    def __init__(self, lib):
        self.lib = lib

    def step(self, dt, bodies, statics, managers, history, fp64=False, anchors=None):
        """Runs one step; returns (bodies', patches[nm], impulses[nm,4],
        partitions[nm] or state64[nb,13]). anchors: an optional
        FRICTION_ANCHORS array of at least nm records that receives the
        step's friction write-back (FP32 only)."""
        b = np.ascontiguousarray(bodies.copy())
        nm = len(managers)
        patches = np.zeros(max(nm, 1), DY_PATCH)
        imp = np.zeros((max(nm, 1), 4), np.float32)
        hist = np.ascontiguousarray(history) if len(history) else np.zeros(1, DY_PATCH)
        st = np.ascontiguousarray(statics)
        mg = np.ascontiguousarray(managers) if nm else np.zeros(1, DY_MANAGER)
        if fp64:
            out = np.zeros((len(b), 13), np.float64)
            rc = self.lib.orc_dy_step64(C.c_float(dt), C.c_uint32(len(b)), P(b), C.c_uint32(len(st)), P(st),
                                        C.c_uint32(nm), P(mg), P(hist), C.c_uint32(len(history)), P(patches),
                                        P(imp), P(out))
        else:
            out = np.zeros(max(nm, 1), np.uint32)
            if anchors is not None:
                assert anchors.flags["C_CONTIGUOUS"] and len(anchors) >= nm
                rc = self.lib.orc_dy_step_anchors(C.c_float(dt), C.c_uint32(len(b)), P(b), C.c_uint32(len(st)),
                                                  P(st), C.c_uint32(nm), P(mg), P(hist), C.c_uint32(len(history)),
                                                  P(patches), P(imp), P(out), P(anchors))
            else:
                rc = self.lib.orc_dy_step(C.c_float(dt), C.c_uint32(len(b)), P(b), C.c_uint32(len(st)), P(st),
                                          C.c_uint32(nm), P(mg), P(hist), C.c_uint32(len(history)), P(patches),
                                          P(imp), P(out))
        if rc != 0:
            raise RuntimeError(f"oracle Dy step failed ({rc})")
        return b, patches[:nm], imp[:nm], out[:nm] if not fp64 else out


def body_state_from(bodies):
    """(n, 13) position, linear velocity, orientation, angular velocity."""
    return np.concatenate([bodies["position"], bodies["linear_velocity"], bodies["orientation"],
                           bodies["angular_velocity"]], axis=1)


class DyScene:
    """Per-step Dy inputs of the reference pile golden (dy_scene.npz),
    regenerated from each step's recorded start state: the oracle SAP and
    PCM (pinned to the reference elsewhere) give the step's pairs and
    manifolds, and the recorded descriptor stream fixes the solver order.
    Friction history and sleep accumulators are carried by the caller (each
    solver carries its own)."""

    _cache = {}

    def __init__(self, orc, g):
        from fyzz_amd.scenes import box_mass_properties
        key = id(g)
        self.g = g
        self.half, self.ids = g["half"], g["ids"]
        self.n = len(self.ids)
        self.steps = int(g["steps"][0])
        self.inv_mass, self.inv_inertia = box_mass_properties(self.half, 1000.0)
        st = np.zeros(1, DY_STATIC)
        st["orientation"][0, 3] = 1.0
        st["static_friction"] = 0.5
        st["dynamic_friction"] = 0.5
        self.statics = st
        if key not in DyScene._cache:
            DyScene._cache[key] = self._regenerate(orc)
        self.managers, self.keys, self.counted = DyScene._cache[key]

    def _regenerate(self, orc):
        g, half, ids, n = self.g, self.half, self.ids, self.n
        pcm = OraclePcm(orc.lib)
        floor = int(g["floor_id"][0])
        slot = {int(v): i for i, v in enumerate(ids)}
        caches = {}
        all_mg, all_keys, all_counted = [], [], []
        for s in range(self.steps):
            st = g["state"][s]
            pos, quat = np.ascontiguousarray(st[:, 0:3]), np.ascontiguousarray(st[:, 6:10])
            b = orc.box_bounds(pos, quat, half, np.full(n, 0.02, np.float32))
            pairs = orc.sap_pairs(b, shape_id=ids, plane_ids=[floor])
            manifolds, new_caches = {}, {}
            for pa, pb in pairs:
                key = (int(pa), int(pb))
                cache = caches.get(key, pcm.new_cache())
                if pa == floor or pb == floor:
                    bx = slot[int(pb if pa == floor else pa)]
                    rec = np.concatenate([half[bx], pos[bx], quat[bx]])
                    cnt, pts, _ = pcm.plane_box(np.zeros(3, np.float32), rec, 0.04, cache)
                    pts = pts.copy()
                    pts[:, 3:6] *= -1.0  # published +y normal
                    bodies_gu = (bx, -1)
                else:
                    # two dynamics: the Gu pair puts the higher body id first
                    f, sc_ = slot[int(pb)], slot[int(pa)]
                    r1 = np.concatenate([half[f], pos[f], quat[f]])
                    r2 = np.concatenate([half[sc_], pos[sc_], quat[sc_]])
                    cnt, pts, _ = pcm.box_box(r1, r2, 0.04, cache)
                    bodies_gu = (f, sc_)
                new_caches[key] = cache
                if cnt:
                    manifolds[key] = (cnt, pts, bodies_gu)
            caches = new_caches
            counted = np.zeros(n, np.uint32)
            for pa, pb in pairs:
                for x in (pa, pb):
                    if int(x) != floor:
                        counted[slot[int(x)]] += 1
            desc = g[f"desc_{s}"]
            mg = np.zeros(len(desc), DY_MANAGER)
            keys = []
            for i, (pa, pb) in enumerate(desc):
                key = (int(pa), int(pb))
                cnt, pts, (b0, b1) = manifolds[key]
                mg[i]["body0"], mg[i]["body1"], mg[i]["count"] = b0, b1, cnt
                for c in range(cnt):
                    mg[i]["contacts"][c]["position"] = pts[c, 0:3]
                    mg[i]["contacts"][c]["normal"] = pts[c, 3:6]
                    mg[i]["contacts"][c]["separation"] = pts[c, 6]
                keys.append(key)
            all_mg.append(mg)
            all_keys.append(keys)
            all_counted.append(counted)
        return all_mg, all_keys, all_counted

    def bodies(self, s, acc_lin, acc_ang):
        g = self.g
        st = g["state"][s]
        b = dy_bodies(self.n)
        b["position"] = st[:, 0:3]
        b["linear_velocity"] = st[:, 3:6]
        b["orientation"] = st[:, 6:10]
        b["angular_velocity"] = st[:, 10:13]
        b["inverse_mass"] = self.inv_mass
        b["inverse_inertia"] = self.inv_inertia
        b["wake_counter"] = g["wake"][s]
        b["flags"] = 4 | (g["ready"][s].astype(np.uint32) << 1)
        b["sleep_linear_acc"] = acc_lin
        b["sleep_angular_acc"] = acc_ang
        b["counted_interactions"] = self.counted[s]
        b["restitution"] = 0.0
        return b

    def step_managers(self, s, history):
        """(managers with history indices, history records) for a
        {pair key: patch} dict of the previous step's patches."""
        mg = self.managers[s].copy()
        hist = []
        for i, key in enumerate(self.keys[s]):
            if key in history:
                mg[i]["history"] = len(hist)
                hist.append(history[key])
            else:
                mg[i]["history"] = 0xFFFFFFFF
        return mg, (np.array(hist, DY_PATCH) if hist else np.zeros(0, DY_PATCH))


def replay_dy_scene(orc, g, on_step, fp64=False, anchors=False):
    """Oracle Dy over the golden pile from synchronized start states; calls
    on_step(s, bodies_out, partitions or FP64 state, managers), with the
    step's friction write-back and manager keys appended when anchors."""
    from fyzz_amd.dy import FRICTION_ANCHORS
    sc = DyScene(orc, g)
    dy = OracleDy(orc.lib)
    history = {}
    acc_lin = np.zeros((sc.n, 3), np.float32)
    acc_ang = np.zeros((sc.n, 3), np.float32)
    for s in range(sc.steps):
        mg, hist = sc.step_managers(s, history)
        an = np.zeros(max(len(mg), 1), FRICTION_ANCHORS) if anchors else None
        out, patches, imp, extra = dy.step(1.0 / 120.0, sc.bodies(s, acc_lin, acc_ang), sc.statics, mg, hist,
                                           fp64=fp64, anchors=an)
        history = {k: patches[i] for i, k in enumerate(sc.keys[s])}
        acc_lin = out["sleep_linear_acc"].copy()
        acc_ang = out["sleep_angular_acc"].copy()
        if anchors:
            on_step(s, out, extra, mg, an[:len(mg)], sc.keys[s])
        else:
            on_step(s, out, extra, mg)


def _components(n, edges):
    """Union-find roots (smallest member) of n nodes over (a, b) edges."""
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    for a, b in edges:
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    return [find(i) for i in range(n)]


def island_pass(b, pairs, prev_pairs, touch_bb, prev_touch_bb, touch_pl, prev_touch_pl):
    """The island pass of a step on body records b (flags: bit0 sleeping,
    bit1 ready, bit3 activating), in place, for dynamic boxes in slots:
    Scene::activate_published_islands (scene.cpp:6148-6262) then
    deactivate_ready_accurate_islands (6465-6536), ordered as Scene::step
    orders them (10747-10793). pairs / prev_pairs: this and the last
    step's broadphase pairs (lo, hi); touch_*: touching box/box pairs and
    plane boxes of this and the last step. Wake propagates through the
    speculative islands (components over both steps' broadphase pairs: the
    pairs lost this step are split only after the wake); sleep takes the
    accurate islands (components over this step's touching pairs) whose
    members are all ready and one awake."""
    f = b["flags"]
    n = len(b)
    prev_set = set(prev_pairs)
    act = set(np.nonzero(f & 8)[0].tolist())
    for lo, hi in pairs:
        if (lo, hi) not in prev_set and not (f[lo] & f[hi] & 1):
            act.update((lo, hi))
    for key in touch_bb:
        if key not in prev_touch_bb:
            act.update(key)
            f[key[0]] &= ~np.uint32(2)
            f[key[1]] &= ~np.uint32(2)
    for i in touch_pl:
        if i not in prev_touch_pl:
            act.add(i)
            f[i] &= ~np.uint32(2)
    if act and (f & 1).any():
        root = _components(n, list(prev_set | set(pairs)))
        wake_roots = {root[i] for i in act}
        for i in range(n):
            if (f[i] & 1) and root[i] in wake_roots:
                f[i] &= ~np.uint32(3)
                b["wake_counter"][i] = max(np.float32(b["wake_counter"][i]), np.float32(0.4))
                b["sleep_linear_acc"][i] = 0.0
                b["sleep_angular_acc"][i] = 0.0
    f &= ~np.uint32(8)
    if ((f & 3) == 2).any():
        root = _components(n, list(touch_bb))
        notready, awake = set(), set()
        for i in range(n):
            if not (f[i] & 2):
                notready.add(root[i])
            if not (f[i] & 1):
                awake.add(root[i])
        for i in range(n):
            r = root[i]
            if not (f[i] & 1) and r not in notready and r in awake:
                f[i] |= np.uint32(1)
                b["wake_counter"][i] = 0.0
                for fld in ("linear_velocity", "angular_velocity", "sleep_linear_acc", "sleep_angular_acc"):
                    b[fld][i] = 0.0
    b["flags"] = f


def island_pass_arrays(b, pairs, prev_pairs, src, touch_bb, prev_touch_bb, touch_pl, prev_touch_pl):
    """island_pass on arrays, for whole piles: pairs / prev_pairs [k, 2]
    (lo, hi) slots; src: each current pair's previous index (-1: found this
    step); touch_bb (aligned with pairs) and prev_touch_bb (aligned with
    prev_pairs) touching flags; touch_pl / prev_touch_pl per box (None: no
    floor). Same semantics; the components come from scipy."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    n = len(b)
    f = b["flags"].astype(np.uint32)

    def labels(edges):
        if len(edges) == 0:
            return np.arange(n)
        m = coo_matrix((np.ones(len(edges), np.int8), (edges[:, 0], edges[:, 1])), shape=(n, n))
        return connected_components(m, directed=False)[1]
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    prev_pairs = np.asarray(prev_pairs, np.int64).reshape(-1, 2)
    act = (f & 8) != 0
    a, c = pairs[:, 0], pairs[:, 1]
    found = (src < 0) & ((f[a] & f[c] & 1) == 0)
    act[a[found]] = True
    act[c[found]] = True
    was = np.zeros(len(pairs), bool)
    if len(prev_pairs):
        ok = src >= 0
        was[ok] = prev_touch_bb[src[ok]] != 0
    new_touch = (touch_bb != 0) & ~was
    for col in (a, c):
        act[col[new_touch]] = True
        f[col[new_touch]] &= ~np.uint32(2)
    if touch_pl is not None:
        nt = (touch_pl != 0) & (prev_touch_pl == 0)
        act[nt] = True
        f[nt] &= ~np.uint32(2)
    if act.any() and (f & 1).any():
        lab = labels(np.concatenate([prev_pairs, pairs]))
        wake = ((f & 1) != 0) & np.isin(lab, np.unique(lab[act]))
        f[wake] &= ~np.uint32(3)
        b["wake_counter"][wake] = np.maximum(b["wake_counter"][wake], np.float32(0.4))
        b["sleep_linear_acc"][wake] = 0.0
        b["sleep_angular_acc"][wake] = 0.0
    f &= ~np.uint32(8)
    if ((f & 3) == 2).any():
        lab = labels(pairs[touch_bb != 0])
        notready = np.unique(lab[(f & 2) == 0])
        awake = np.unique(lab[(f & 1) == 0])
        sleep = ((f & 1) == 0) & ~np.isin(lab, notready) & np.isin(lab, awake)
        f[sleep] |= np.uint32(1)
        b["wake_counter"][sleep] = 0.0
        for fld in ("linear_velocity", "angular_velocity", "sleep_linear_acc", "sleep_angular_acc"):
            b[fld][sleep] = 0.0
    b["flags"] = f


class OracleWorld:
    """The whole step pipeline composed from the oracle's SAP, PCM and Dy,
    in the device world's manager order (touching box/box managers in pair
    order (lo slot, hi slot), then touching plane managers by slot); friction
    history is kept for managers that touched on the previous step, box/box
    caches while the pair stays in the broadphase list, plane caches per box."""

    def __init__(self, orc, scene):
        from fyzz_amd.scenes import box_mass_properties
        self.orc = orc
        self.pcm = OraclePcm(orc.lib)
        self.dy = OracleDy(orc.lib)
        n = scene.n
        self.n = n
        self.half = scene.half
        inv_mass, inv_inertia = box_mass_properties(scene.half, scene.density)
        b = dy_bodies(n)
        b["position"] = scene.pos
        q = scene.quat.astype(np.float32)
        ln = np.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
        b["orientation"] = q / ln[:, None]
        b["linear_velocity"] = scene.vel
        b["angular_velocity"] = scene.ang
        b["inverse_mass"] = inv_mass
        b["inverse_inertia"] = inv_inertia
        b["static_friction"] = scene.friction
        b["dynamic_friction"] = scene.friction
        b["restitution"] = scene.restitution
        self.bodies = b
        self.floor = scene.floor
        st = np.zeros(1, DY_STATIC)
        st["orientation"][0, 3] = 1.0
        st["position"][0, 1] = scene.floor_height
        st["static_friction"] = scene.friction
        st["dynamic_friction"] = scene.friction
        st["restitution"] = scene.restitution
        self.statics = st
        self.bb_caches = {}
        self.pl_caches = [self.pcm.new_cache() for _ in range(n)]
        self.history = {}
        self.last = None
        self.prev_pairs, self.prev_touch_bb, self.prev_touch_pl = [], set(), set()

    def step(self, dt):
        n, half, b = self.n, self.half, self.bodies
        pos = np.ascontiguousarray(b["position"])
        quat = np.ascontiguousarray(b["orientation"])
        bounds = self.orc.box_bounds(pos, quat, half, np.full(n, 0.02, np.float32))
        pairs = self.orc.sap_pairs(bounds)
        order = np.lexsort((pairs[:, 1], pairs[:, 0]))
        pairs = pairs[order]
        counted = (np.bincount(pairs.ravel(), minlength=n) + (1 if self.floor else 0)).astype(np.uint32)
        rec = lambda i: np.concatenate([half[i], pos[i], quat[i]])  # noqa: E731
        mgrs, keys = [], []
        caches = {}
        for lo, hi in pairs:
            key = ("bb", int(lo), int(hi))
            cache = self.bb_caches.get(key, self.pcm.new_cache())
            cnt, pts, _ = self.pcm.box_box(rec(hi), rec(lo), 0.04, cache)
            caches[key] = cache
            if cnt:
                mgrs.append((int(hi), int(lo), cnt, pts))
                keys.append(key)
        self.bb_caches = caches
        if self.floor:
            fp = np.ascontiguousarray(self.statics["position"][0])
            for i in range(n):
                cnt, pts, _ = self.pcm.plane_box(fp, rec(i), 0.04, self.pl_caches[i])
                if cnt:
                    pts = pts.copy()
                    pts[:, 3:6] *= -1.0
                    mgrs.append((i, -1, cnt, pts))
                    keys.append(("pl", i))
        # the island pass: wakes, then sleeps; sleeping islands leave the
        # solver list (their manifolds stay as they are)
        cur_pairs = [(int(lo), int(hi)) for lo, hi in pairs]
        touch_bb = {(k[1], k[2]) for k in keys if k[0] == "bb"}
        touch_pl = {k[1] for k in keys if k[0] == "pl"}
        island_pass(self.bodies, cur_pairs, self.prev_pairs, touch_bb, self.prev_touch_bb, touch_pl,
                    self.prev_touch_pl)
        self.prev_pairs, self.prev_touch_bb, self.prev_touch_pl = cur_pairs, touch_bb, touch_pl
        b = self.bodies
        keep = [k for k, m in enumerate(mgrs) if not (b["flags"][m[0]] & 1)]
        mgrs = [mgrs[k] for k in keep]
        keys = [keys[k] for k in keep]
        mg = np.zeros(len(mgrs), DY_MANAGER)
        hist = []
        for k, (b0, b1, cnt, pts) in enumerate(mgrs):
            mg[k]["body0"], mg[k]["body1"], mg[k]["count"] = b0, b1, cnt
            for c in range(cnt):
                mg[k]["contacts"][c]["position"] = pts[c, 0:3]
                mg[k]["contacts"][c]["normal"] = pts[c, 3:6]
                mg[k]["contacts"][c]["separation"] = pts[c, 6]
            if keys[k] in self.history:
                mg[k]["history"] = len(hist)
                hist.append(self.history[keys[k]])
            else:
                mg[k]["history"] = 0xFFFFFFFF
        hist = np.array(hist, DY_PATCH) if hist else np.zeros(0, DY_PATCH)
        b = b.copy()
        b["counted_interactions"] = counted
        out, patches, imp, parts = self.dy.step(dt, b, self.statics, mg, hist)
        self.last_dy_inputs = (b, mg, hist)
        self.history = {k: patches[i] for i, k in enumerate(keys)}
        self.bodies = out
        self.last = (pairs, mg, imp, parts)
        self.last_keys = keys
        return out

    # ---- authored edits between steps, as the reference's Scene setters
    # make them (scene.cpp:589-781, 1022-1147); i is a box slot
    def _wake(self, i, queue=False):
        """The body's own wake: wake counter raised to 0.4 s, sleep flags
        cleared; a sleeping body also drops its sleep filter and is queued
        for the next step's island wake (bit3), as is a peer (queue)."""
        b = self.bodies
        if b["flags"][i] & 1:
            b["sleep_linear_acc"][i] = 0.0
            b["sleep_angular_acc"][i] = 0.0
            queue = True
        b["wake_counter"][i] = max(np.float32(b["wake_counter"][i]), np.float32(0.4))
        b["flags"][i] = (np.uint32(b["flags"][i]) & np.uint32(~3 & 0xFFFFFFFF)) | np.uint32(8 if queue else 0)

    def set_pose(self, i, position, orientation):
        """Scene::set_pose: normalized orientation (a translation-only edit
        passes the stored one back unchanged), the body's wake, every
        current pair's PCM cache reset; when the pose changed, the pairs'
        friction history erased and touching peers that sleep or are ready
        to sleep woken (scene.cpp:589-700)."""
        b = self.bodies
        p = np.asarray(position, np.float32)
        q = np.asarray(orientation, np.float32)
        mag = np.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
        nq = (q / mag).astype(np.float32)
        if np.array_equal(q, b["orientation"][i]):
            nq = b["orientation"][i].copy()
        changed = not (np.array_equal(p, b["position"][i]) and np.array_equal(nq, b["orientation"][i]))
        self._wake(i)
        for key in list(self.bb_caches):
            if i in key[1:]:
                self.bb_caches[key] = self.pcm.new_cache()
        self.pl_caches[i] = self.pcm.new_cache()
        if changed:
            for key in list(self.history):
                if (key[0] == "bb" and i in key[1:]) or key == ("pl", i):
                    del self.history[key]
            # peers of every touching manifold (sleeping islands' too)
            for lo, hi in self.prev_touch_bb:
                other = hi if lo == i else lo if hi == i else -1
                if other >= 0 and (b["flags"][other] & 3):
                    self._wake(other, queue=True)
        b["position"][i] = p
        b["orientation"][i] = nq

    def set_velocity(self, i, linear=None, angular=None):
        """Scene::set_velocity / set_angular_velocity: the value and the
        body's wake (scene.cpp:730-781)."""
        if linear is not None:
            self.bodies["linear_velocity"][i] = np.asarray(linear, np.float32)
        if angular is not None:
            self.bodies["angular_velocity"][i] = np.asarray(angular, np.float32)
        self._wake(i)

    def set_material(self, i, static_friction, dynamic_friction, restitution):
        b = self.bodies
        b["static_friction"][i] = static_friction
        b["dynamic_friction"][i] = dynamic_friction
        b["restitution"][i] = restitution

    def set_dynamics(self, i, linear_damping, angular_damping, max_linear, max_angular, gravity,
                     position_iterations=4, velocity_iterations=1):
        b = self.bodies
        b["linear_damping"][i] = linear_damping
        b["angular_damping"][i] = angular_damping
        b["max_linear_velocity"][i] = max_linear
        b["max_angular_velocity"][i] = max_angular
        f = int(b["flags"][i])
        b["flags"][i] = (f | 4) if gravity else (f & ~4)
        b["iterations"][i] = position_iterations | (velocity_iterations << 8)

    def shape_pair(self, key):
        """Canonical shape-id pair of a manager key (floor = body 0, box
        slot s = body s + 1, as the world tests author them)."""
        if key[0] == "bb":
            return (key[1] + 1, key[2] + 1)
        return (0, key[1] + 1)

    def body_states(self):
        return body_state_from(self.bodies)


def world_step_events(lib, ow, prev):
    """The contact events of an OracleWorld's last step: its solver
    managers (awake islands) and the touches that started or ended. prev:
    the sorted touching shape pairs of the step before (the value this
    returns for it). Returns (events, offsets, impulses, points, point
    impulses, touching shape pairs). A touch that persists inside a
    sleeping island is not reported, and is not lost either."""
    pairs, mg, imp, _ = ow.last
    keys = [ow.shape_pair(k) for k in ow.last_keys]
    touching = sorted([ow.shape_pair(("bb",) + k) for k in ow.prev_touch_bb] +
                      [ow.shape_pair(("pl", i)) for i in ow.prev_touch_pl])
    tset, kset = set(touching), set(keys)
    prev = [p for p in prev if not (p in tset and p not in kset)]
    out = oracle_contact_events(lib, mg, imp, keys, [True] * len(keys), prev)
    return out + (touching,)


def oracle_contact_events(lib, managers, impulses, keys, reversed_flags, prev_pairs, body_of_shape=None):
    """orc_contact_events over touching managers (any order; sorted here
    into Gu pair order). managers: DY_MANAGER records with published
    points; impulses [n, 4] per Gu point; keys: canonical (shape a, shape b)
    per manager; reversed_flags per manager; prev_pairs: sorted (a, b)
    touching last step. Returns (events [k, 6] u32, offsets u64, impulses
    f64, points [m, 7] f32, point impulses f64)."""
    n = len(keys)
    order = sorted(range(n), key=lambda i: keys[i])
    cur = np.zeros((max(n, 1), 36), np.uint32)
    imp = np.zeros((max(n, 1), 4), np.float32)
    for r, i in enumerate(order):
        a, b = keys[i]
        ba = a if body_of_shape is None else body_of_shape[a]
        bb = b if body_of_shape is None else body_of_shape[b]
        cur[r, :6] = [a, b, ba, bb, int(reversed_flags[i]), int(managers[i]["count"])]
        pts = np.zeros((4, 7), np.float32)
        for k in range(int(managers[i]["count"])):
            c = managers[i]["contacts"][k]
            pts[k, 0:3] = c["position"]
            pts[k, 3:6] = c["normal"]
            pts[k, 6] = c["separation"]
        cur[r, 6:34] = pts.reshape(-1).view(np.uint32)
        imp[r] = impulses[i]
    prev = np.ascontiguousarray(np.asarray(prev_pairs, np.uint32).reshape(-1, 2)) if len(prev_pairs) else \
        np.zeros((1, 2), np.uint32)
    cap = n + len(prev_pairs) + 1
    ev = np.zeros((cap, 6), np.uint32)
    off = np.zeros(cap, np.uint64)
    eimp = np.zeros(cap, np.float64)
    pt = np.zeros((4 * n + 1, 7), np.float32)
    pimp = np.zeros(4 * n + 1, np.float64)
    ne, npt = C.c_uint32(), C.c_uint32()
    rc = lib.orc_contact_events(C.c_uint32(n), P(cur), P(imp), C.c_uint32(len(prev_pairs)), P(prev), C.c_uint32(cap),
                                P(ev), P(off), P(eimp), C.c_uint32(len(pt)), P(pt), P(pimp), C.byref(ne),
                                C.byref(npt))
    assert rc == 0
    k, m = ne.value, npt.value
    return ev[:k], off[:k], eimp[:k], pt[:m], pimp[:m]


class OracleDiff:
    """The frozen-row rollout of the differentiable form in FP64
    (oracle/orc_dy.cpp orc_diff_*): built from the initial body records and
    each step's recorded managers and history. nominal_fp32 freezes the
    rows along the FP32 restatement's trajectory (the one an FP32
    implementation freezes), else along the FP64 twin's."""

    def __init__(self, lib, bodies0, statics, managers_per_step, history_per_step, dt, nominal_fp32=False):
        self.lib = lib
        lib.orc_diff_create.restype = C.c_void_p
        self.nb = len(bodies0)
        nm = np.array([len(m) for m in managers_per_step], np.uint32)
        nh = np.array([len(h) for h in history_per_step], np.uint32)
        mg = np.concatenate([np.asarray(m, DY_MANAGER) for m in managers_per_step]) if nm.sum() else \
            np.zeros(1, DY_MANAGER)
        hs = [np.asarray(h, DY_PATCH) for h in history_per_step if len(h)]
        hist = np.concatenate(hs) if hs else np.zeros(1, DY_PATCH)
        b0 = np.ascontiguousarray(bodies0)
        st = np.ascontiguousarray(statics)
        self.h = lib.orc_diff_create(C.c_uint32(self.nb), P(b0), C.c_uint32(len(st)), P(st), C.c_uint32(len(nm)),
                                     P(nm), P(np.ascontiguousarray(mg)), P(nh), P(np.ascontiguousarray(hist)),
                                     C.c_float(dt), C.c_int(1 if nominal_fp32 else 0))
        if not self.h:
            raise RuntimeError("oracle nominal rollout failed")
        self.h = C.c_void_p(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_diff_destroy(self.h)

    def eval(self, v0, w0, mu):
        out = np.zeros((self.nb, 3))
        self.lib.orc_diff_eval(self.h, P(np.ascontiguousarray(v0, np.float64)), P(np.ascontiguousarray(w0, np.float64)),
                               P(np.ascontiguousarray(mu, np.float64)), P(out))
        return out

    def grad(self, v0, w0, mu, adj_p):
        gv, gw, gm, pf = np.zeros((self.nb, 3)), np.zeros((self.nb, 3)), np.zeros(self.nb), np.zeros((self.nb, 3))
        self.lib.orc_diff_grad(self.h, P(np.ascontiguousarray(v0, np.float64)), P(np.ascontiguousarray(w0, np.float64)),
                               P(np.ascontiguousarray(mu, np.float64)), P(np.ascontiguousarray(adj_p, np.float64)),
                               P(gv), P(gw), P(gm), P(pf))
        return gv, gw, gm, pf


def _keys64(pairs):
    p = np.asarray(pairs, np.uint64).reshape(-1, 2)
    return (p[:, 0] << np.uint64(32)) | p[:, 1]


class StepOracle:
    """One whole world step composed from the oracle's SAP, PCM (batched
    C++ entries) and Dy, started from a device world's synchronized state:
    its body records, its live pair list with the box/box manifold caches
    aligned to it, its plane caches and its friction history (the world's
    persistent_state()). The manager order is the world's (touching
    box/box managers in pair order, then touching plane managers by slot).
    scene_first (optional) splits the boxes into independent scenes:
    pairs of boxes in different scenes are dropped, as the batched world's
    per-scene broadphase never forms them. Sized for whole piles: the pair
    work runs in C++ on `threads` host threads."""

    def __init__(self, orc, half, floor_height=0.0, floor_friction=0.5, floor_restitution=0.0, scene_first=None,
                 threads=8):
        self.orc = orc
        self.lib = orc.lib
        self.half = np.ascontiguousarray(half, np.float32)
        self.n = len(self.half)
        st = np.zeros(1, DY_STATIC)
        st["orientation"][0, 3] = 1.0
        st["position"][0, 1] = floor_height
        st["static_friction"] = floor_friction
        st["dynamic_friction"] = floor_friction
        st["restitution"] = floor_restitution
        self.statics = st
        self.scene_first = None if scene_first is None else np.asarray(scene_first, np.int64)
        self.threads = threads
        self.dy = OracleDy(orc.lib)

    def pairs(self, bodies):
        pos = np.ascontiguousarray(bodies["position"])
        quat = np.ascontiguousarray(bodies["orientation"])
        off = np.full(self.n, 0.02, np.float32)
        if self.scene_first is None:
            pairs = self.orc.sap_pairs(self.orc.box_bounds(pos, quat, self.half, off))
        else:
            # one sweep per scene (scenes overlap in space but never pair)
            parts = []
            for a, b in zip(self.scene_first[:-1], self.scene_first[1:]):
                if b - a < 2:
                    continue
                bounds = self.orc.box_bounds(pos[a:b], quat[a:b], self.half[a:b], off[a:b])
                parts.append(self.orc.sap_pairs(bounds).astype(np.uint32) + np.uint32(a))
            pairs = np.concatenate(parts) if parts else np.zeros((0, 2), np.uint32)
        return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]

    def step(self, bodies, prev_pairs, state, dt):
        """Returns a dict: pairs, found, lost, counts_bb, pts_bb, counts_pl,
        pts_pl, managers, bodies (after the step), impulses, partitions."""
        n = self.n
        pairs = self.pairs(bodies)
        found, lost = self.orc.found_lost(prev_pairs, pairs)
        k_cur, k_prev = _keys64(pairs), _keys64(prev_pairs)
        j = np.searchsorted(k_prev, k_cur)
        jc = np.minimum(j, max(len(k_prev) - 1, 0))
        kept = (j < len(k_prev)) & (k_prev[jc] == k_cur) if len(k_prev) else np.zeros(len(k_cur), bool)
        src = np.where(kept, jc, -1)
        caches = np.zeros((max(len(pairs), 1), 56), np.float32)
        if kept.any():
            caches[: len(pairs)][kept] = state["caches_bb"][jc[kept]]
        pos = np.ascontiguousarray(bodies["position"])
        quat = np.ascontiguousarray(bodies["orientation"])
        first = np.ascontiguousarray(pairs[:, 1], np.uint32)  # Gu order: the higher slot first
        second = np.ascontiguousarray(pairs[:, 0], np.uint32)
        npairs = len(pairs)
        pts_bb = np.zeros((max(npairs, 1), 4, 7), np.float32)
        cnt_bb = np.zeros(max(npairs, 1), np.uint32)
        fl_bb = np.zeros(max(npairs, 1), np.uint32)
        rc = self.lib.orc_pcm_box_box_many(C.c_uint32(npairs), P(first), P(second), P(pos), P(quat), P(self.half),
                                           C.c_float(0.04), P(caches), P(pts_bb), P(cnt_bb), P(fl_bb),
                                           C.c_uint32(self.threads))
        if rc != 0:
            raise OverflowError(f"oracle PCM box/box manifold invalid at pair {-1 - rc}")
        cpl = np.ascontiguousarray(state["caches_pl"].copy())
        pts_pl = np.zeros((n, 4, 7), np.float32)
        cnt_pl = np.zeros(n, np.uint32)
        fl_pl = np.zeros(n, np.uint32)
        fp = np.ascontiguousarray(self.statics["position"][0])
        rc = self.lib.orc_pcm_plane_box_many(C.c_uint32(n), P(fp), P(pos), P(quat), P(self.half), C.c_float(0.04),
                                             P(cpl), P(pts_pl), P(cnt_pl), P(fl_pl), C.c_uint32(self.threads))
        if rc != 0:
            raise OverflowError(f"oracle PCM plane/box manifold invalid at box {-1 - rc}")
        pts_pl[:, :, 3:6] *= -1.0  # the scene publishes the plane's +y normal
        # the island pass (wakes, then sleeps) on the step's bodies
        b = bodies.copy()
        island_pass_arrays(b, pairs, prev_pairs, src, cnt_bb[:npairs], state["touch_bb"], cnt_pl,
                           state["touch_pl"])
        asleep = (b["flags"] & 1) != 0
        # touching managers of awake islands in the world's order, with
        # their history slots
        tb = np.nonzero(cnt_bb[:npairs])[0]
        tb = tb[~asleep[first[tb]]]
        tp = np.nonzero(cnt_pl)[0]
        tp = tp[~asleep[tp]]
        nm = len(tb) + len(tp)
        mg = np.zeros(nm, DY_MANAGER)
        mg["body0"][: len(tb)] = first[tb]
        mg["body1"][: len(tb)] = second[tb]
        mg["count"][: len(tb)] = cnt_bb[tb]
        mg["body0"][len(tb):] = tp
        mg["body1"][len(tb):] = -1
        mg["count"][len(tb):] = cnt_pl[tp]
        pts = np.concatenate([pts_bb[tb], pts_pl[tp]]) if nm else np.zeros((0, 4, 7), np.float32)
        live = np.arange(4)[None, :] < mg["count"][:, None]
        pts = np.where(live[:, :, None], pts, 0.0).astype(np.float32)
        mg["contacts"]["position"] = pts[:, :, 0:3]
        mg["contacts"]["normal"] = pts[:, :, 3:6]
        mg["contacts"]["separation"] = pts[:, :, 6]
        # history: box/box slot = previous pair index, plane slot = P_prev + box
        hslot = np.concatenate([src[tb], len(k_prev) + tp]).astype(np.int64)
        step_no = state["step"]
        ok = (hslot >= 0) & (step_no > 0)
        ok[ok] &= state["stamps"][hslot[ok]] == step_no
        hist_idx = np.full(nm, 0xFFFFFFFF, np.uint32)
        hist_idx[ok] = np.arange(int(ok.sum()), dtype=np.uint32)
        mg["history"] = hist_idx
        hist = state["history"][hslot[ok]] if ok.any() else np.zeros(0, DY_PATCH)
        counted = np.bincount(pairs.ravel(), minlength=n).astype(np.uint32) + 1
        b["counted_interactions"] = counted
        from fyzz_amd.dy import FRICTION_ANCHORS
        an = np.zeros(max(nm, 1), FRICTION_ANCHORS)
        out, patches, imp, parts = self.dy.step(dt, b, self.statics, mg, hist, anchors=an)
        return {"pairs": pairs, "found": found, "lost": lost, "counts_bb": cnt_bb[:npairs],
                "pts_bb": pts_bb[:npairs], "counts_pl": cnt_pl, "pts_pl": pts_pl, "managers": mg, "bodies": out,
                "impulses": imp, "partitions": parts, "patches": patches, "anchors": an[:nm]}
