"""Write the island sleep/wake golden from the compiled reference (build
container only).

sleep_scene.npz: the 16-box layout of author_scene.npz (six two-box
stacks and four single boxes resting on the floor) stepped 150 times at
1/120 s by the reference NativeCpu scene. The boxes settle, their wake
counters run out and whole islands (a stack, a single box) go to sleep
(Scene::deactivate_ready_accurate_islands, scene.cpp:6465-6536). Then:
  before step 90   set_pose lifts single 3 above sleeping single 0: it
                   falls onto it, the broadphase finds the pair and the
                   touch wakes single 0's island
                   (Scene::activate_published_islands, 6148-6262)
  before step 100  set_velocity on sleeping single 1 (an authored wake)
  before step 110  set_pose on stack 4's top, sleeping: the pose edit wakes
                   it and its island
Each body stays in at most one dynamic/dynamic contact, so the manager
order cannot change a result. Per step s: state[s] (13 floats per box at
the start of step s), sleeping[s], ready[s], wake[s]; the edits as data
(edits, the codes of author_scene.npz); ev_{s} [k, 6] u32 (body0, body1,
kind, contact count, shape0, shape1) and evimp_{s} f64, the contact events
of step s (notifications on: a touch that persists inside a sleeping
island is not reported).

    python tests/golden/make_sleep_golden.py
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "fyzzxyzzy-sdk_amd"))
sys.path.insert(0, str(Path(__file__).resolve().parent))
from make_author_golden import author_scene  # noqa: E402
from make_sap_golden import RefScene, load_ref, P  # noqa: E402

STEPS = 150


def main(path: Path, events: bool = True):
    ref = load_ref()
    ref.ref_contact_events.restype = C.c_size_t
    sc = RefScene(ref)
    if events:
        sc.ok(ref.ref_set_contact_notifications(sc.s, 1))
    ev_rec = {}
    floor = sc.add_floor(0.0)
    sc.ok(ref.ref_set_material(sc.s, C.c_uint32(floor), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0)))
    half, pos, quat = author_scene()
    n = len(half)
    ids = [sc.add_box(half[b], pos[b]) for b in range(n)]
    for b in range(n):
        sc.set_pose(ids[b], pos[b], quat[b])
        sc.ok(ref.ref_set_material(sc.s, C.c_uint32(ids[b]), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0)))
    st = np.zeros(14, np.float32)

    def state_of(b):
        sc.ok(ref.ref_body_state(sc.s, C.c_uint32(ids[b]), P(st)))
        return st.copy()

    edits = []  # (step, code, box, values)
    states, sleeping, readies, wakes = [], [], [], []
    for s in range(STEPS):
        if s == 90:
            top = state_of(12)  # single 0
            p = np.array([top[0], top[1] + half[12][1] + half[15][1] + 0.15, top[2]], np.float32)
            q = np.array([0.0, 0.0, 0.0, 1.0], np.float32)
            sc.set_pose(ids[15], p, q)
            edits.append((s, 0, 15, list(p) + list(q)))
        if s == 100:
            v = np.array([0.6, 0.0, -0.2], np.float32)
            sc.ok(ref.ref_set_velocity(sc.s, C.c_uint32(ids[13]), P(v)))
            edits.append((s, 1, 13, list(v)))
        if s == 110:
            cur = state_of(9)
            p = (cur[:3] + np.array([0.0, 0.002, 0.0], np.float32)).astype(np.float32)
            sc.set_pose(ids[9], p, cur[6:10].copy())
            edits.append((s, 0, 9, list(p) + list(cur[6:10])))
        snap = np.zeros((n, 13), np.float32)
        sl = np.zeros(n, np.uint8)
        wk = np.zeros(n, np.float32)
        rd = np.zeros(n, np.uint8)
        for b in range(n):
            v = state_of(b)
            snap[b] = v[:13]
            sl[b] = v[13] > 0
            w = C.c_float()
            r = C.c_int()
            sc.ok(ref.ref_sleep_state(sc.s, C.c_uint32(ids[b]), C.byref(w), C.byref(r)))
            wk[b] = w.value
            rd[b] = r.value
        states.append(snap)
        sleeping.append(sl)
        readies.append(rd)
        wakes.append(wk)
        sc.ok(ref.ref_step(sc.s, C.c_float(1.0 / 120.0)))
        if events:
            npts = C.c_size_t()
            ne = ref.ref_contact_events(sc.s, None, None, None, 0, None, None, 0, C.byref(npts))
            ev = np.zeros((max(ne, 1), 6), np.uint32)
            off = np.zeros(max(ne, 1), np.uint64)
            imp = np.zeros(max(ne, 1), np.float64)
            pt = np.zeros((max(npts.value, 1), 7), np.float32)
            pimp = np.zeros(max(npts.value, 1), np.float64)
            ref.ref_contact_events(sc.s, P(ev), P(off), P(imp), ne, P(pt), P(pimp), npts.value, C.byref(npts))
            ev_rec[f"ev_{s}"] = ev[:ne]
            ev_rec[f"evimp_{s}"] = imp[:ne]
    snap = np.array([state_of(b)[:13] for b in range(n)], np.float32)
    states.append(snap)
    sleeping = np.array(sleeping)
    ed = np.zeros((len(edits), 11), np.float64)
    for i, (s, code, b, vals) in enumerate(edits):
        ed[i, :3] = (s, code, b)
        ed[i, 3:3 + len(vals)] = vals
    np.savez_compressed(path, half=half, pos=pos, quat=quat, ids=np.array(ids, np.uint32),
                        state=np.array(states), sleeping=sleeping, ready=np.array(readies), wake=np.array(wakes),
                        edits=ed, steps=np.array([STEPS]), **ev_rec)
    first = {b: int(np.argmax(sleeping[:, b])) if sleeping[:, b].any() else -1 for b in range(n)}
    print(path.name, "boxes", n, "steps", STEPS, "first sleeping step per box", first)
    print("sleeping count per 10 steps", sleeping.sum(1)[::10].tolist())


if __name__ == "__main__":
    main(Path(__file__).resolve().parent / "sleep_scene.npz")
