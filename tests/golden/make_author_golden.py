"""Write the mid-run authoring golden from the compiled reference (build
container only).

author_scene.npz: 16 boxes on the floor (six two-box stacks and four
single boxes, seeded sizes and yaws, friction 0.5, restitution 0) stepped
45 times at 1/120 s by the reference NativeCpu scene, with authored edits
between steps (each body takes part in at most one dynamic/dynamic
contact, so the solver order cannot change a result):
  before step 20  set_pose on stack 0's top (shifted and turned), a
                  translation-only set_pose on stack 1's top (the stored
                  orientation passed back), set_velocity on single 0,
                  set_angular_velocity on single 1, set_material on stack
                  2's bottom, set_dynamics on stack 3's top (damping) and
                  on single 2 (gravity off, with an upward velocity)
  before step 25  a 1 mm translation of stack 5's top and an idempotent
                  set_pose of stack 4's top (its own pose passed back)
  before step 30  set_pose on single 3 (lifted and moved: it falls again)
Per step s: state[s] (13 floats per box at the start of step s: position,
linear velocity, orientation, angular velocity), wake[s], ready[s]; the
edits are stored as data (edit_*).

    python tests/golden/make_author_golden.py
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "fyzzxyzzy-sdk_amd"))
sys.path.insert(0, str(Path(__file__).resolve().parent))
from fyzz_amd import scenes  # noqa: E402
from make_sap_golden import RefScene, load_ref, P  # noqa: E402

STEPS = 45
SEED = scenes.SEED_BASE + 31


def author_scene():
    """Half extents, resting positions and yaw quaternions of the 16 boxes
    (stacks first: bottom then top; then the singles)."""
    half, pos = [], []
    hs = 0.3 + 0.2 * scenes.uniform(SEED, 0, 48).reshape(16, 3)
    k = 0
    for s in range(6):
        x, z = 3.0 * (s % 3) - 3.0, 3.0 * (s // 3) - 1.5
        hb, ht = hs[k], hs[k + 1]
        yb = hb[1] + 0.005
        yt = 2 * hb[1] + ht[1] + 0.01
        half += [hb, ht]
        pos += [(x, yb, z), (x + 0.05, yt, z - 0.04)]
        k += 2
    for s in range(4):
        h = hs[k]
        half.append(h)
        pos.append((3.0 * s - 4.5, h[1] + 0.005, 4.5))
        k += 1
    half = np.array(half, np.float32)
    pos = np.array(pos, np.float32)
    quat = scenes.normalize_quat32(scenes.yaw_quat(scenes.uniform(SEED, 1, 16) * 0.5 * np.pi))
    return half, pos, quat


EDITS = {
    # step: [(kind, box index, values...)]
    20: [("pose", 1, (0.10, 0.0, -0.05), (0.0, 0.0998334, 0.0, 0.9950042)),
         ("position", 3, (-0.02, 0.0, 0.06)),
         ("velocity", 12, (1.0, 0.0, 0.3)),
         ("angular", 13, (0.0, 2.0, 0.5)),
         ("material", 4, (0.9, 0.8, 0.0)),
         ("dynamics", 7, (0.8, 0.3, 1.0e16, 100.0), 1),
         ("dynamics", 14, (0.0, 0.05, 1.0e16, 100.0), 0),
         ("velocity", 14, (0.0, 0.5, 0.0))],
    25: [("position", 11, (0.001, 0.0, 0.0)),
         ("position", 9, (0.0, 0.0, 0.0))],
    30: [("pose", 15, (0.0, 0.9, 7.0), (0.0, 0.0, 0.0, 1.0))],
}


def main(path: Path):
    ref = load_ref()
    sc = RefScene(ref)
    floor = sc.add_floor(0.0)
    sc.ok(ref.ref_set_material(sc.s, C.c_uint32(floor), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0)))
    half, pos, quat = author_scene()
    n = len(half)
    ids = [sc.add_box(half[b], pos[b]) for b in range(n)]
    for b in range(n):
        sc.set_pose(ids[b], pos[b], quat[b])
        sc.ok(ref.ref_set_material(sc.s, C.c_uint32(ids[b]), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0)))
    st = np.zeros(14, np.float32)
    states, wakes, readies = [], [], []
    rec = {"half": half, "pos": pos, "quat": quat, "ids": np.array(ids, np.uint32)}
    applied = []
    for s in range(STEPS):
        for e in EDITS.get(s, []):
            kind, b = e[0], ids[e[1]]
            if kind == "pose":
                sc.set_pose(b, np.array(e[2], np.float32), np.array(e[3], np.float32))
            elif kind == "position":
                sc.ok(ref.ref_body_state(sc.s, C.c_uint32(b), P(st)))
                p = (st[:3] + np.array(e[2], np.float32)).astype(np.float32)
                sc.set_pose(b, p, st[6:10].copy())
                e = ("pose", e[1], tuple(p.tolist()), tuple(st[6:10].tolist()))
            elif kind == "velocity":
                sc.ok(ref.ref_set_velocity(sc.s, C.c_uint32(b), P(np.array(e[2], np.float32))))
            elif kind == "angular":
                sc.ok(ref.ref_set_angular_velocity(sc.s, C.c_uint32(b), P(np.array(e[2], np.float32))))
            elif kind == "material":
                sc.ok(ref.ref_set_material(sc.s, C.c_uint32(b), *[C.c_float(v) for v in e[2]]))
            elif kind == "dynamics":
                sc.ok(ref.ref_set_dynamics(sc.s, C.c_uint32(b), P(np.array(e[2], np.float32)), C.c_int(e[3]),
                                           C.c_int(4), C.c_int(1)))
            applied.append((s, e))
        snap = np.zeros((n, 13), np.float32)
        wk = np.zeros(n, np.float32)
        rd = np.zeros(n, np.uint8)
        for b in range(n):
            sc.ok(ref.ref_body_state(sc.s, C.c_uint32(ids[b]), P(st)))
            assert st[13] == 0.0, "a body went to sleep: the scene must stay awake"
            snap[b] = st[:13]
            w = C.c_float()
            r = C.c_int()
            sc.ok(ref.ref_sleep_state(sc.s, C.c_uint32(ids[b]), C.byref(w), C.byref(r)))
            wk[b] = w.value
            rd[b] = r.value
        states.append(snap)
        wakes.append(wk)
        readies.append(rd)
        sc.ok(ref.ref_step(sc.s, C.c_float(1.0 / 120.0)))
    snap = np.zeros((n, 13), np.float32)
    for b in range(n):
        sc.ok(ref.ref_body_state(sc.s, C.c_uint32(ids[b]), P(st)))
        snap[b] = st[:13]
    states.append(snap)
    rec["state"] = np.array(states)
    rec["wake"] = np.array(wakes)
    rec["ready"] = np.array(readies)
    # edits as data: step, kind code, box, 8 values (pose: p3 q4; position
    # edits are stored as the pose they passed; velocity / angular: v3;
    # material: sf df rest; dynamics: 4 floats + gravity)
    codes = {"pose": 0, "velocity": 1, "angular": 2, "material": 3, "dynamics": 4}
    ed = np.zeros((len(applied), 11), np.float64)
    for i, (s, e) in enumerate(applied):
        vals = list(e[2]) + (list(e[3]) if e[0] == "pose" else [e[3]] if e[0] == "dynamics" else [])
        ed[i, :3] = (s, codes[e[0]], e[1])
        ed[i, 3:3 + len(vals)] = vals
    rec["edits"] = ed
    rec["steps"] = np.array([STEPS])
    np.savez_compressed(path, **rec)
    print(path.name, "boxes", n, "steps", STEPS, "edits", len(applied))


if __name__ == "__main__":
    main(Path(__file__).resolve().parent / "author_scene.npz")
