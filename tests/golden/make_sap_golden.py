"""Write SAP golden vectors from the compiled reference NativeCpu scene.

Runs only in the build container, where oracle/_ref/libfyzzref.so is built
from the reference's own sources (`make -C oracle ref`). Writes
tests/golden/sap_*.npz: the scene inputs and, per frame, the reference's
sorted endpoints (3 axes), live pairs, found and lost lists.

    python tests/golden/make_sap_golden.py
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "fyzzxyzzy-sdk_amd"))
from fyzz_amd import scenes  # noqa: E402

P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731


def load_ref():
    ref = C.CDLL(str(ROOT / "oracle" / "_ref" / "libfyzzref.so"))
    ref.ref_scene_create.restype = C.c_void_p
    ref.ref_last_error.restype = C.c_char_p
    for f in ["ref_pairs", "ref_found", "ref_lost", "ref_endpoints"]:
        getattr(ref, f).restype = C.c_size_t
    return ref


class RefScene:
    def __init__(self, ref):
        self.ref = ref
        self.s = C.c_void_p(ref.ref_scene_create())

    def ok(self, rc):
        if rc != 0:
            raise RuntimeError(self.ref.ref_last_error().decode())

    def add_floor(self, h=0.0):
        i = C.c_uint32()
        self.ok(self.ref.ref_add_floor(self.s, C.c_float(h), C.byref(i)))
        return i.value

    def add_box(self, half, pos, density=1000.0):
        i = C.c_uint32()
        self.ok(self.ref.ref_add_dynamic_box(self.s, P(np.ascontiguousarray(half, np.float32)),
                                             P(np.ascontiguousarray(pos, np.float32)), C.c_float(density),
                                             C.byref(i)))
        return i.value

    def set_pose(self, i, pos, q):
        self.ok(self.ref.ref_set_pose(self.s, C.c_uint32(i), P(np.ascontiguousarray(pos, np.float32)),
                                      P(np.ascontiguousarray(q, np.float32))))

    def set_position(self, i, pos):
        self.ok(self.ref.ref_set_position(self.s, C.c_uint32(i), P(np.ascontiguousarray(pos, np.float32))))

    def advance(self):
        self.ok(self.ref.ref_advance_collision(self.s, 1))

    def pairlist(self, fn):
        m = fn(self.s, None, 0)
        out = np.zeros(2 * m, np.uint32)
        fn(self.s, P(out), m)
        return out.reshape(-1, 2)

    def endpoints(self):
        vals, ids, mx = [], [], []
        for a in range(3):
            k = self.ref.ref_endpoints(self.s, C.c_uint(a), None, None, None, 0)
            v = np.zeros(k, np.uint32)
            d = np.zeros(k, np.uint32)
            m = np.zeros(k, np.uint8)
            self.ref.ref_endpoints(self.s, C.c_uint(a), P(v), P(d), P(m), k)
            vals.append(v)
            ids.append(d)
            mx.append(m)
        return np.stack(vals), np.stack(ids), np.stack(mx)


def sap_frames(path: Path, n=300, frames=4, rotated=False, floor=False, seed=scenes.SEED_BASE + 11,
               density=2.5):
    ref = load_ref()
    # a cube dense enough to give a few hundred pairs
    half, pos = scenes.sap_scene(n, seed=seed, extent=density * n ** (1 / 3))
    quat = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    if rotated:
        quat = scenes.normalize_quat32(scenes.random_quat(scenes.uniform(seed, 20, n), scenes.uniform(seed, 21, n),
                                                          scenes.uniform(seed, 22, n)))
    sc = RefScene(ref)
    first = sc.add_floor(-1000.0) if floor else None
    ids = np.array([sc.add_box(half[i], pos[i]) for i in range(n)], np.uint32)
    if rotated:
        for i in range(n):
            sc.set_pose(int(ids[i]), pos[i], quat[i])
        # keep the orientation as the scene stores it (set_pose renormalises,
        # scene.cpp:589-600): the stage inputs are the stored poses
        st = np.zeros(14, np.float32)
        for i in range(n):
            sc.ok(ref.ref_body_state(sc.s, C.c_uint32(int(ids[i])), P(st)))
            quat[i] = st[6:10]
    rec = {"half": half, "pos0": pos.copy(), "quat": quat, "ids": ids,
           "floor_id": np.array([-1 if first is None else first], np.int64)}
    cur = pos.copy()
    for f in range(frames):
        if f > 0:
            slots, delta = scenes.sap_moves(n, f, seed=seed, fraction=0.25)
            delta = delta * 20.0  # larger moves so that pairs change
            cur[slots] = (cur[slots].astype(np.float64) + delta).astype(np.float32)
            for s in slots:
                if rotated:
                    sc.set_pose(int(ids[s]), cur[s], quat[s])
                    st = np.zeros(14, np.float32)
                    sc.ok(ref.ref_body_state(sc.s, C.c_uint32(int(ids[s])), P(st)))
                    if not np.array_equal(st[6:10], quat[s]):
                        raise RuntimeError("stored orientation changed on re-posing")
                else:
                    sc.set_position(int(ids[s]), cur[s])
            rec[f"moved_{f}"] = slots.astype(np.uint32)
            rec[f"pos_{f}"] = cur[slots].copy()
        sc.advance()
        v, d, m = sc.endpoints()
        rec[f"ep_value_{f}"] = v
        rec[f"ep_id_{f}"] = d
        rec[f"ep_max_{f}"] = m
        rec[f"pairs_{f}"] = sc.pairlist(ref.ref_pairs)
        rec[f"found_{f}"] = sc.pairlist(ref.ref_found)
        rec[f"lost_{f}"] = sc.pairlist(ref.ref_lost)
        print(path.name, "frame", f, "pairs", len(rec[f"pairs_{f}"]), "found", len(rec[f"found_{f}"]),
              "lost", len(rec[f"lost_{f}"]))
    rec["frames"] = np.array([frames])
    np.savez_compressed(path, **rec)


if __name__ == "__main__":
    out = Path(__file__).resolve().parent
    sap_frames(out / "sap_aabb.npz")
    sap_frames(out / "sap_rotated_floor.npz", n=240, rotated=True, floor=True, seed=scenes.SEED_BASE + 12,
               density=4.0)
