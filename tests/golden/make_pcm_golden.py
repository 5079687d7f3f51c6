"""Write PCM golden vectors from the compiled reference (build container only).

pcm_stateless.npz: random box/box pose pairs with the reference's
  source_pcm_box_box_candidates, source_pcm_box_batch_selection and
  source_pcm_box_box_gjk_epa outputs.
pcm_scene.npz: a small falling pile stepped by the reference NativeCpu
  scene: the poses at the start of every step, and for every 4th step the
  live pairs and published manifolds (pair, count, 4 x position/normal/
  separation) that the step's narrowphase produced from those poses.

    python tests/golden/make_pcm_golden.py
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "fyzzxyzzy-sdk_amd"))
sys.path.insert(0, str(Path(__file__).resolve().parent))
from fyzz_amd import scenes  # noqa: E402
from make_sap_golden import RefScene, load_ref, P  # noqa: E402


def box_record(h, p, q):
    return np.concatenate([h, p, q]).astype(np.float32)


def random_pairs(n, seed):
    """Pose pairs near contact: half extents U[0.2,1.0], second box placed
    at a random direction with a distance around the sum of radii."""
    u = lambda s, k=n: scenes.uniform(seed, s, k)  # noqa: E731
    ha = np.stack([0.2 + 0.8 * u(a) for a in range(3)], 1)
    hb = np.stack([0.2 + 0.8 * u(3 + a) for a in range(3)], 1)
    qa = scenes.normalize_quat32(scenes.random_quat(u(6), u(7), u(8)))
    qb = scenes.normalize_quat32(scenes.random_quat(u(9), u(10), u(11)))
    # a third of the pairs keep axis-aligned or shared orientations (face contacts)
    qa[0::3] = np.array([0, 0, 0, 1], np.float32)
    qb[0::3] = np.array([0, 0, 0, 1], np.float32)
    qb[1::6] = qa[1::6]
    d = np.stack([scenes.normal(seed, 12 + a, n) for a in range(3)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ra = np.linalg.norm(ha, axis=1)
    rb = np.linalg.norm(hb, axis=1)
    dist = (0.5 + 0.6 * u(15)) * (ra + rb)
    pa = np.stack([(u(16 + a) - 0.5) * 20 for a in range(3)], 1)
    pb = pa + d * dist[:, None]
    out = []
    for i in range(n):
        out.append((box_record(ha[i], pa[i], qa[i]), box_record(hb[i], pb[i], qb[i])))
    return out


def stateless(path: Path, n=1500, seed=scenes.SEED_BASE + 21):
    ref = load_ref()
    A, B, CNT, OK, CANDS, SEL, GJK_OK, GJK = [], [], [], [], [], [], [], []
    cd = 0.04
    for a, b in random_pairs(n, seed):
        cand = np.zeros(64 * 10, np.float32)
        cnt = C.c_uint32()
        ok = C.c_int()
        if ref.ref_pcm_box_box_candidates(P(a), P(b), C.c_float(cd), P(cand), 64, C.byref(cnt), C.byref(ok)) != 0:
            raise RuntimeError("reference candidates failed")
        sel = np.full(4, 255, np.uint8)
        if cnt.value > 4:
            ref.ref_pcm_batch_selection(P(cand), cnt.value, C.c_float(1.0), P(sel))
        g = np.zeros(10, np.float32)
        gok = C.c_int()
        ref.ref_pcm_gjk_epa(P(a), P(b), C.c_float(cd), P(g), C.byref(gok))
        A.append(a)
        B.append(b)
        CNT.append(cnt.value)
        OK.append(ok.value)
        CANDS.append(cand.reshape(64, 10)[:16])
        SEL.append(sel)
        GJK_OK.append(gok.value)
        GJK.append(g)
    CNT = np.array(CNT, np.uint32)
    print(path.name, "pairs", n, "overlapping", int(np.sum(OK)), "max candidates", CNT.max(),
          "zero-witness overlaps", int(np.sum((CNT == 0) & (np.array(OK) == 1))), "gjk contacts", int(np.sum(GJK_OK)))
    np.savez_compressed(path, a=np.array(A), b=np.array(B), count=CNT, overlapping=np.array(OK, np.uint8),
                        cands=np.array(CANDS), sel=np.array(SEL), gjk_ok=np.array(GJK_OK, np.uint8),
                        gjk=np.array(GJK), contact_distance=np.array([cd], np.float32))


def scene(path: Path, side=3, steps=72, every=4, seed=scenes.SEED_BASE + 22):
    """side^3 boxes of mixed sizes on a grid over the floor, seeded yaw and
    velocity, stepped at 1/120 s."""
    ref = load_ref()
    sc = RefScene(ref)
    floor = sc.add_floor(0.0)
    n = side ** 3
    idx = np.arange(n)
    i, k, j = idx % side, (idx // side) % side, idx // (side * side)
    half = np.stack([0.3 + 0.3 * scenes.uniform(seed, a, n) for a in range(3)], 1).astype(np.float32)
    pos = np.stack([(i - (side - 1) / 2) * 1.3, 0.7 + j * 1.3, (k - (side - 1) / 2) * 1.3], 1).astype(np.float32)
    ids = np.array([sc.add_box(half[b], pos[b]) for b in range(n)], np.uint32)
    quat = scenes.normalize_quat32(scenes.random_quat(scenes.uniform(seed, 3, n), scenes.uniform(seed, 4, n),
                                                      scenes.uniform(seed, 5, n)))
    vel = np.stack([scenes.normal(seed, 6 + a, n) * 0.5 for a in range(3)], 1).astype(np.float32)
    for b in range(n):
        sc.set_pose(int(ids[b]), pos[b], quat[b])
        sc.ok(ref.ref_set_velocity(sc.s, C.c_uint32(int(ids[b])), P(vel[b])))
        sc.ok(ref.ref_set_material(sc.s, C.c_uint32(int(ids[b])), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0)))
    rec = {"half": half, "ids": ids, "floor_id": np.array([floor], np.int64)}
    st = np.zeros(14, np.float32)
    poses = []
    ref.ref_manifolds.restype = C.c_size_t
    for s in range(steps):
        pose = np.zeros((n, 7), np.float32)
        for b in range(n):
            sc.ok(ref.ref_body_state(sc.s, C.c_uint32(int(ids[b])), P(st)))
            pose[b, :3] = st[0:3]
            pose[b, 3:] = st[6:10]
        poses.append(pose)
        sc.ok(ref.ref_step(sc.s, C.c_float(1.0 / 120.0)))
        if s % every == every - 1:
            pairs = sc.pairlist(ref.ref_pairs)
            m = ref.ref_manifolds(sc.s, None, None, None, 0)
            mp = np.zeros(2 * m, np.uint32)
            mc = np.zeros(m, np.uint8)
            pts = np.zeros(m * 4 * 7, np.float32)
            ref.ref_manifolds(sc.s, P(mp), P(mc), P(pts), m)
            rec[f"pairs_{s}"] = pairs
            rec[f"mpairs_{s}"] = mp.reshape(-1, 2)
            rec[f"mcount_{s}"] = mc
            rec[f"mpoints_{s}"] = pts.reshape(m, 4, 7)
    rec["poses"] = np.array(poses)
    rec["steps"] = np.array([steps])
    rec["every"] = np.array([every])
    print(path.name, "boxes", n, "steps", steps, "manifolds at last", len(rec[f"mcount_{steps - 1}"]))
    np.savez_compressed(path, **rec)


if __name__ == "__main__":
    out = Path(__file__).resolve().parent
    stateless(out / "pcm_stateless.npz")
    scene(out / "pcm_scene.npz")
