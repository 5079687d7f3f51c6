"""GPU: the HIP PCM narrowphase against the oracle and the reference goldens.

Tolerance (contact points, normals, separations): 1e-5 relative with a
1e-6 absolute floor for values near zero. Counts, overlap flags, batch
selections and regeneration/GJK decisions must match exactly. The kernels
keep the reference's operation order, so most values are bit-identical;
the test reports how many are.
"""
from pathlib import Path

import numpy as np
import pytest

from fyzz_amd import scenes
from oracle_bind import OraclePcm

GOLDEN = Path(__file__).resolve().parent / "golden"
pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-5, 1e-6


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def close(got, want, what):
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=ATOL, err_msg=what)


def test_candidates_and_selection_match_reference(ctx, torch_cuda):
    from fyzz_amd import pcm
    torch = torch_cuda
    g = np.load(GOLDEN / "pcm_stateless.npz")
    cd = float(g["contact_distance"][0])
    out, counts, ov, sel = pcm.candidates(ctx, dev(torch, g["a"]), dev(torch, g["b"]), cd)
    torch.cuda.synchronize()
    out, counts, ov, sel = out.cpu().numpy(), counts.cpu().numpy(), ov.cpu().numpy(), sel.cpu().numpy()
    np.testing.assert_array_equal(ov, g["overlapping"])
    np.testing.assert_array_equal(counts, g["count"])
    exact = 0
    for i in range(len(counts)):
        c = counts[i]
        close(out[i, :c], g["cands"][i, :c], f"pair {i}")
        exact += np.array_equal(out[i, :c], g["cands"][i, :c])
        if c > 4:
            np.testing.assert_array_equal(sel[i], g["sel"][i], err_msg=f"pair {i}")
    print(f"candidate sets bit-identical: {exact}/{len(counts)}")


def test_gjk_epa_matches_reference(ctx, torch_cuda):
    from fyzz_amd import pcm
    torch = torch_cuda
    g = np.load(GOLDEN / "pcm_stateless.npz")
    cd = float(g["contact_distance"][0])
    out, ok = pcm.gjk_epa(ctx, dev(torch, g["a"]), dev(torch, g["b"]), cd)
    torch.cuda.synchronize()
    out, ok = out.cpu().numpy(), ok.cpu().numpy()
    np.testing.assert_array_equal(ok, g["gjk_ok"])
    m = ok.astype(bool)
    close(out[m], g["gjk"][m], "gjk witnesses")


class PairCaches:
    """Host bookkeeping of persistent cache records keyed by pair (the
    reference keys them by the contact manager's EdgeIndex)."""

    def __init__(self, capacity, torch):
        from fyzz_amd import pcm
        self.dev = pcm.new_caches(capacity, "cuda")
        self.host = {}
        self.slot_of = {}
        self.free = list(range(capacity))[::-1]
        self.torch = torch

    def slots(self, keys, orc_pcm):
        live = set(keys)
        for k in list(self.slot_of):
            if k not in live:
                self.free.append(self.slot_of.pop(k))
                del self.host[k]
        out = []
        for k in keys:
            if k not in self.slot_of:
                s = self.free.pop()
                self.slot_of[k] = s
                self.dev[s].zero_()
                self.host[k] = orc_pcm.new_cache()
            out.append(self.slot_of[k])
        return np.array(out, np.int32)


def _compare_frame(orc_pcm, pos, quat, half, bb_pairs, plane_boxes, caches, ctx, torch, tag):
    """Run the GPU and the oracle for one frame; return #bit-identical contacts."""
    from fyzz_amd import pcm
    keys = [("bb",) + tuple(p) for p in bb_pairs] + [("pl", int(b)) for b in plane_boxes]
    slots = caches.slots(keys, orc_pcm)
    nb = len(bb_pairs)
    dpos, dquat, dhalf = dev(torch, pos), dev(torch, quat), dev(torch, half)
    exact = total = 0
    results = []
    if nb:
        first = dev(torch, np.array([p[0] for p in bb_pairs], np.int32))
        second = dev(torch, np.array([p[1] for p in bb_pairs], np.int32))
        results.append(("bb", pcm.box_box(ctx, first, second, dpos, dquat, dhalf, caches.dev,
                                          cache_index=dev(torch, slots[:nb]))))
    if len(plane_boxes):
        results.append(("pl", pcm.plane_box(ctx, dev(torch, np.asarray(plane_boxes, np.int32)), dpos, dquat, dhalf,
                                            caches.dev, cache_index=dev(torch, slots[nb:]))))
    torch.cuda.synchronize()
    for kind, (c, n, f) in results:
        c, n, f = c.cpu().numpy(), n.cpu().numpy(), f.cpu().numpy()
        items = bb_pairs if kind == "bb" else plane_boxes
        for i, it in enumerate(items):
            if kind == "bb":
                key = ("bb",) + tuple(it)
                r1 = np.concatenate([half[it[0]], pos[it[0]], quat[it[0]]])
                r2 = np.concatenate([half[it[1]], pos[it[1]], quat[it[1]]])
                wn, wp, wf = orc_pcm.box_box(r1, r2, 0.04, caches.host[key])
            else:
                key = ("pl", int(it))
                rec = np.concatenate([half[it], pos[it], quat[it]])
                wn, wp, wf = orc_pcm.plane_box(np.zeros(3, np.float32), rec, 0.04, caches.host[key])
                wp = wp.copy()
                wp[:, 3:6] *= -1.0  # published +y normal
            assert n[i] == wn, f"{tag} {key}: {n[i]} vs {wn} contacts"
            assert (f[i] & 3) == wf, f"{tag} {key}: flags {f[i]} vs {wf}"
            close(c[i, :wn], wp, f"{tag} {key}")
            exact += int(np.array_equal(c[i, :wn], wp))
            total += 1
    gpu_cache = caches.dev.cpu().numpy()
    for key, s in caches.slot_of.items():
        close(gpu_cache[s][:55], caches.host[key][:55], f"{tag} cache {key}")
        assert gpu_cache[s][55].view(np.uint32) == caches.host[key][55].view(np.uint32), f"{tag} cache {key}"
    return exact, total


def test_scene_manifolds_match_oracle_and_reference(ctx, orc, torch_cuda):
    """Replays the reference pile's 72 steps: per step, GPU manifolds and
    caches == the oracle (whose manifolds equal the reference's)."""
    torch = torch_cuda
    orc_pcm = OraclePcm(orc.lib)
    g = np.load(GOLDEN / "pcm_scene.npz")
    half, ids = g["half"], g["ids"]
    floor = int(g["floor_id"][0])
    n = len(ids)
    caches = PairCaches(4096, torch)
    exact = total = 0
    for s in range(int(g["steps"][0])):
        pose = g["poses"][s]
        pos, quat = np.ascontiguousarray(pose[:, :3]), np.ascontiguousarray(pose[:, 3:])
        b = orc.box_bounds(pos, quat, half, np.full(n, 0.02, np.float32))
        pairs = orc.sap_pairs(b, shape_id=ids, plane_ids=[floor])
        slot = {int(v): i for i, v in enumerate(ids)}
        bb = [(slot[int(p[1])], slot[int(p[0])]) for p in pairs if p[0] != floor and p[1] != floor]
        pl = [slot[int(p[1] if p[0] == floor else p[0])] for p in pairs if p[0] == floor or p[1] == floor]
        e, t = _compare_frame(orc_pcm, pos, quat, half, bb, pl, caches, ctx, torch, f"step {s}")
        exact += e
        total += t
    print(f"bit-identical manifolds: {exact}/{total}")


def test_random_walk_manifolds_match_oracle(ctx, orc, torch_cuda):
    """1,500 independent box pairs near contact, 12 frames of small seeded
    pose perturbations: cache refresh, translation/rotation thresholds,
    regeneration, GJK/EPA continuation and the five-point reduction all run."""
    torch = torch_cuda
    orc_pcm = OraclePcm(orc.lib)
    g = np.load(GOLDEN / "pcm_stateless.npz")
    a, b = g["a"], g["b"]
    m = len(a)
    half = np.concatenate([a[:, :3], b[:, :3]]).astype(np.float32)
    pos = np.concatenate([a[:, 3:6], b[:, 3:6]]).astype(np.float32)
    quat = np.concatenate([a[:, 6:], b[:, 6:]]).astype(np.float32)
    bb = [(i, m + i) for i in range(m)]
    caches = PairCaches(2 * m, torch)
    seed = scenes.SEED_BASE + 40
    exact = total = 0
    for f in range(12):
        if f:
            k = 2 * m
            pos = (pos + np.stack([scenes.normal(seed, 10 * f + c, k) for c in range(3)], 1) * 0.01).astype(np.float32)
            w = np.stack([scenes.normal(seed, 10 * f + 3 + c, k) for c in range(3)], 1) * 0.02
            dq = np.concatenate([0.5 * w, np.ones((k, 1))], 1)
            qd = quat.astype(np.float64)
            # q' = normalize(dq * q)
            x1, y1, z1, w1 = dq.T
            x2, y2, z2, w2 = qd.T
            qn = np.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 + y1 * w2 + z1 * x2 - x1 * z2,
                           w1 * z2 + z1 * w2 + x1 * y2 - y1 * x2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
            quat = scenes.normalize_quat32(qn)
            # every fourth frame leaves half of the boxes where they are
            if f % 4 == 0:
                pos[::2] = prev_pos[::2]
                quat[::2] = prev_quat[::2]
        prev_pos, prev_quat = pos.copy(), quat.copy()
        e, t = _compare_frame(orc_pcm, pos, quat, half, bb, [], caches, ctx, torch, f"frame {f}")
        exact += e
        total += t
    print(f"bit-identical manifolds: {exact}/{total}")
