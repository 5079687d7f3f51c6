"""Write Dy golden vectors from the compiled reference (build container only).

dy_scene.npz: a 27-box pile (mixed sizes, seeded orientation and velocity,
friction 0.5, restitution 0 on boxes and floor) stepped 72 times at 1/120 s
by the reference NativeCpu scene. Per step s:
  state[s]      body state at the start of step s: position, linear
                velocity, orientation, angular velocity (13 floats),
  wake[s], ready[s]  wake counter and ready-for-sleep flag,
  desc_{s}      the step's Dy descriptor stream in solver-task order, as the
                shape pair of each descriptor's manifold (Gu order: the
                pair's canonical (lo, hi) shape ids),
  part_{s}      each descriptor's conflict partition (global),
  task_{s}      the tasks' cumulative partition ends,
  fr_{s}, frpos_{s}, frimp_{s}  the friction write-back of the step
                (Scene::solved_friction_patches): per patch the shape pair,
                anchor count and broken flag, anchor positions and world
                impulses (f64).
Contacts are not stored: the oracle's narrowphase reproduces the
reference's manifolds bit-exactly (tests/test_pcm_oracle.py).

dy_events.npz: the same run with contact notifications on (the simulation
is unchanged; the script checks the states against dy_scene.npz). Per step
s: ev_{s} [n, 6] u32 (body0, body1, kind 0 found / 1 persists / 2 lost,
contact count, shape0, shape1), evoff_{s} u64 point offsets, evimp_{s} f64
summed normal impulses, pt_{s} [m, 7] f32 (position, outward normal,
separation), ptimp_{s} f64 per-point normal impulses.

    python tests/golden/make_dy_golden.py
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


def dy_scene(path: Path, side=3, steps=72, seed=scenes.SEED_BASE + 23, spacing=1.15, base=0.62,
             events_path: Path | None = None):
    ref = load_ref()
    for f in ("ref_dy_descriptors", "ref_dy_tasks", "ref_manifolds", "ref_contact_events", "ref_friction_patches"):
        getattr(ref, f).restype = C.c_size_t
    sc = RefScene(ref)
    if events_path is not None:
        sc.ok(ref.ref_set_contact_notifications(sc.s, 1))
    events = {}
    floor = sc.add_floor(0.0)
    sc.ok(ref.ref_set_material(sc.s, C.c_uint32(floor), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0)))
    n = side ** 3
    idx = np.arange(n)
    i, k, j = idx % side, (idx // side) % side, idx // (side * side)
    half = np.stack([0.3 + 0.25 * scenes.uniform(seed, a, n) for a in range(3)], 1).astype(np.float32)
    pos = np.stack([(i - (side - 1) / 2) * spacing, base + j * spacing, (k - (side - 1) / 2) * spacing], 1)
    pos = pos.astype(np.float32)
    ids = np.array([sc.add_box(half[b], pos[b]) for b in range(n)], np.uint32)
    yaw = scenes.uniform(seed, 3, n) * 2 * np.pi
    quat = scenes.normalize_quat32(scenes.yaw_quat(yaw))
    vel = np.stack([scenes.normal(seed, 6 + a, n) * 0.3 for a in range(3)], 1).astype(np.float32)
    for b in range(n):
        sc.set_pose(int(ids[b]), pos[b], quat[b])
        sc.ok(ref.ref_set_velocity(sc.s, C.c_uint32(int(ids[b])), P(vel[b])))
        sc.ok(ref.ref_set_material(sc.s, C.c_uint32(int(ids[b])), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0)))
    st = np.zeros(14, np.float32)
    states, wakes, readies = [], [], []
    rec = {"half": half, "ids": ids, "floor_id": np.array([floor], np.int64)}

    def snapshot():
        s = np.zeros((n, 13), np.float32)
        wk = np.zeros(n, np.float32)
        rd = np.zeros(n, np.uint8)
        for b in range(n):
            sc.ok(ref.ref_body_state(sc.s, C.c_uint32(int(ids[b])), P(st)))
            s[b] = st[:13]
            w = C.c_float()
            r = C.c_int()
            sc.ok(ref.ref_sleep_state(sc.s, C.c_uint32(int(ids[b])), C.byref(w), C.byref(r)))
            wk[b] = w.value
            rd[b] = r.value
        states.append(s)
        wakes.append(wk)
        readies.append(rd)

    for s in range(steps):
        snapshot()
        sc.ok(ref.ref_step(sc.s, C.c_float(1.0 / 120.0)))
        m = ref.ref_manifolds(sc.s, None, None, None, 0)
        mp = np.zeros(2 * m, np.uint32)
        mc = np.zeros(m, np.uint8)
        pts = np.zeros(m * 4 * 7, np.float32)
        ref.ref_manifolds(sc.s, P(mp), P(mc), P(pts), m)
        mp = mp.reshape(-1, 2)
        d = ref.ref_dy_descriptors(sc.s, None, None, None, None, 0)
        di = np.zeros(d, np.uint32)
        b0 = np.zeros(d, np.uint32)
        b1 = np.zeros(d, np.uint32)
        pt = np.zeros(d, np.uint64)
        ref.ref_dy_descriptors(sc.s, P(di), P(b0), P(b1), P(pt), d)
        t = ref.ref_dy_tasks(sc.s, None, None, 0)
        me = np.zeros(t, np.uint64)
        pe = np.zeros(t, np.uint64)
        ref.ref_dy_tasks(sc.s, P(me), P(pe), t)
        if events_path is not None:
            npts = C.c_size_t()
            ne = ref.ref_contact_events(sc.s, None, None, None, 0, None, None, 0, C.byref(npts))
            ev = np.zeros((max(ne, 1), 6), np.uint32)
            off = np.zeros(max(ne, 1), np.uint64)
            imp = np.zeros(max(ne, 1), np.float64)
            pt = np.zeros((max(npts.value, 1), 7), np.float32)
            pimp = np.zeros(max(npts.value, 1), np.float64)
            ref.ref_contact_events(sc.s, P(ev), P(off), P(imp), ne, P(pt), P(pimp), npts.value, C.byref(npts))
            events[f"ev_{s}"] = ev[:ne]
            events[f"evoff_{s}"] = off[:ne]
            events[f"evimp_{s}"] = imp[:ne]
            events[f"pt_{s}"] = pt[:npts.value]
            events[f"ptimp_{s}"] = pimp[:npts.value]
        nf = ref.ref_friction_patches(sc.s, None, None, None, 0)
        fp = np.zeros((max(nf, 1), 4), np.uint32)
        fpos = np.zeros((max(nf, 1), 2, 3), np.float32)
        fimp = np.zeros((max(nf, 1), 2, 3), np.float64)
        ref.ref_friction_patches(sc.s, P(fp), P(fpos), P(fimp), nf)
        rec[f"fr_{s}"] = fp[:nf]
        rec[f"frpos_{s}"] = fpos[:nf]
        rec[f"frimp_{s}"] = fimp[:nf]
        rec[f"desc_{s}"] = mp[di]
        rec[f"body_{s}"] = np.stack([b0, b1], 1)
        rec[f"part_{s}"] = pt
        rec[f"task_{s}"] = pe
    snapshot()
    rec["state"] = np.array(states)
    rec["wake"] = np.array(wakes)
    rec["ready"] = np.array(readies)
    rec["steps"] = np.array([steps])
    print(path.name, "boxes", n, "steps", steps, "descriptors at last", len(rec[f"desc_{steps - 1}"]),
          "max partitions", max(int(rec[f"part_{s}"].max()) + 1 if len(rec[f"part_{s}"]) else 0
                                for s in range(steps)))
    if events_path is not None:
        events["steps"] = np.array([steps])
        np.savez_compressed(events_path, **events)
        print(events_path.name, "events", sum(len(v) for k, v in events.items() if k.startswith("ev_")))
        return rec
    np.savez_compressed(path, **rec)
    return rec


if __name__ == "__main__":
    here = Path(__file__).resolve().parent
    plain = dy_scene(here / "dy_scene.npz")
    with_events = dy_scene(here / "dy_scene.npz", events_path=here / "dy_events.npz")
    assert np.array_equal(plain["state"], with_events["state"]), "notifications changed the simulation"
