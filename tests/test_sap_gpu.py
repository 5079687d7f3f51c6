"""GPU: the HIP SAP stages against the oracle and the reference goldens.

Bit-exact bar: encoded bounds, sorted endpoints (3 axes), live pairs and
found/lost lists must equal the oracle / golden arrays exactly.
"""
from pathlib import Path

import numpy as np
import pytest

from fyzz_amd import scenes

GOLDEN = Path(__file__).resolve().parent / "golden"
pytestmark = pytest.mark.gpu


def dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def random_boxes(seed, n, extent_scale=2.0, rotated=True):
    half, pos = scenes.sap_scene(n, seed=seed, extent=extent_scale * n ** (1 / 3))
    if rotated:
        quat = scenes.normalize_quat32(scenes.random_quat(scenes.uniform(seed, 20, n), scenes.uniform(seed, 21, n),
                                                          scenes.uniform(seed, 22, n)))
        quat[::5] = np.array([0, 0, 0, 1], np.float32)  # identity fast path
    else:
        quat = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    return half, pos, quat, np.full(n, 0.02, np.float32)


@pytest.mark.parametrize("n,seed", [(1, 3), (257, 4), (5000, 5), (70000, 6)])
def test_bounds_and_sort_match_oracle(ctx, orc, torch_cuda, n, seed):
    from fyzz_amd import sap
    torch = torch_cuda
    half, pos, quat, off = random_boxes(seed, n)
    # equal keys: duplicate some boxes exactly so ties are broken by slot
    if n > 10:
        pos[n // 2:n // 2 + 5] = pos[0]
        half[n // 2:n // 2 + 5] = half[0]
        quat[n // 2:n // 2 + 5] = quat[0]
    want_b = orc.box_bounds(pos, quat, half, off)
    b, err = sap.box_bounds(ctx, dev(torch, pos), dev(torch, quat), dev(torch, half), dev(torch, off))
    vals, pays = sap.sort_endpoints(ctx, b)
    torch.cuda.synchronize()
    assert int(err.item()) == 0
    np.testing.assert_array_equal(u32(b), want_b)
    wv, wp = orc.sort_endpoints(want_b)
    np.testing.assert_array_equal(u32(vals), wv)
    np.testing.assert_array_equal(u32(pays), wp)


def test_bounds_flags_nonfinite(ctx, torch_cuda):
    from fyzz_amd import sap
    torch = torch_cuda
    pos = np.array([[0, 0, 0], [3e38, 0, 0]], np.float32)
    quat = np.tile(np.array([0, 0, 0, 1], np.float32), (2, 1))
    half = np.full((2, 3), 3e38, np.float32)
    _, err = sap.box_bounds(ctx, dev(torch, pos), dev(torch, quat), dev(torch, half),
                            dev(torch, np.zeros(2, np.float32)))
    torch.cuda.synchronize()
    assert int(err.item()) & 1


@pytest.mark.parametrize("name", ["sap_aabb.npz", "sap_rotated_floor.npz"])
def test_incremental_sap_matches_reference_golden(ctx, torch_cuda, name):
    """Per frame: sorted endpoints, live pairs, found, lost == the compiled
    reference NativeCpu scene (plane pairs are the reference's Cartesian
    product and are not SAP output, so they are excluded here)."""
    from fyzz_amd import sap
    torch = torch_cuda
    g = np.load(GOLDEN / name)
    half, quat, ids = g["half"], g["quat"], g["ids"]
    n = len(ids)
    floor = int(g["floor_id"][0])
    pos = g["pos0"].copy()
    off = np.full(n, 0.02, np.float32)
    bp = sap.SapBroadphase(ctx, ids, np.ones(n, np.uint8))
    dq, dh, do = dev(torch, quat), dev(torch, half), dev(torch, off)

    def box_only(p):
        return p[(p[:, 0] != floor) & (p[:, 1] != floor)] if floor >= 0 else p

    for f in range(int(g["frames"][0])):
        dirty = None
        if f:
            pos[g[f"moved_{f}"]] = g[f"pos_{f}"]
            dirty = dev(torch, g[f"moved_{f}"].astype(np.int32))
        bp.advance(dev(torch, pos), dq, dh, do, dirty=dirty)
        vals, pays = bp.endpoints()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(u32(vals), g[f"ep_value_{f}"])
        np.testing.assert_array_equal(ids[u32(pays) >> 1], g[f"ep_id_{f}"])
        for which, key in [(0, "pairs"), (1, "found"), (2, "lost")]:
            got = ids[bp.fetch(which)]
            np.testing.assert_array_equal(got, box_only(g[f"{key}_{f}"]), err_msg=f"frame {f} {key}")
    bp.close()


def _run_frames(ctx, orc, torch, n, frames, seed, scale, move_scale, fraction, rotated, check_endpoints=True,
                delta_capacity=None):
    from fyzz_amd import sap
    half, pos, quat, off = random_boxes(seed, n, extent_scale=scale, rotated=rotated)
    body = np.arange(n, dtype=np.uint32)
    body[1::11] = body[0::11][: len(body[1::11])]
    dyn = (scenes.uniform(seed, 31, n) < 0.8).astype(np.uint8)
    bp = sap.SapBroadphase(ctx, body, dyn, delta_capacity=delta_capacity)
    dq, dh, do = dev(torch, quat), dev(torch, half), dev(torch, off)
    prev = np.zeros((0, 2), np.uint32)
    for f in range(frames):
        dirty = None
        if f:
            slots, delta = scenes.sap_moves(n, f, seed=seed, fraction=fraction)
            pos[slots] = (pos[slots].astype(np.float64) + delta * move_scale).astype(np.float32)
            # a few listed boxes do not move at all; a few are listed twice
            dl = np.concatenate([slots, slots[:7], np.arange(3)]).astype(np.int32)
            dirty = dev(torch, dl)
        bp.advance(dev(torch, pos), dq, dh, do, dirty=dirty)
        b = orc.box_bounds(pos, quat, half, off)
        want = orc.sap_pairs(b, body_id=body, dynamic=dyn)
        wf, wl = orc.found_lost(prev, want)
        np.testing.assert_array_equal(bp.fetch(0), want, err_msg=f"pairs frame {f}")
        np.testing.assert_array_equal(bp.fetch(1), wf, err_msg=f"found frame {f}")
        np.testing.assert_array_equal(bp.fetch(2), wl, err_msg=f"lost frame {f}")
        if check_endpoints:
            vals, pays = bp.endpoints()
            wv, wp = orc.sort_endpoints(b)
            np.testing.assert_array_equal(u32(vals), wv)
            np.testing.assert_array_equal(u32(pays), wp)
        prev = want
    np.testing.assert_equal(bp.counts()[3], 0)
    bp.close()


def test_incremental_sap_concentrated_changes(ctx, orc, torch_cuda):
    """Four huge low-slot boxes jump across the scene each frame, so tens of
    thousands of pair changes share one lower slot range: the change list
    is larger than the one-workgroup sort and one key-range bucket holds
    more than it can sort in LDS, so the chunk-sort + merge rounds run."""
    from fyzz_amd import sap
    torch = torch_cuda
    n = 12000
    half, pos, quat, off = random_boxes(21, n, extent_scale=1.5, rotated=False)
    lo, hi = pos.min(0), pos.max(0)
    span = hi - lo
    half[:4] = (0.45 * span).astype(np.float32)
    corners = [lo + 0.3 * span, hi - 0.3 * span]
    body = np.arange(n, dtype=np.uint32)
    dyn = np.ones(n, np.uint8)
    bp = sap.SapBroadphase(ctx, body, dyn, delta_capacity=2_000_000)
    dq, dh, do = dev(torch, quat), dev(torch, half), dev(torch, off)
    prev = np.zeros((0, 2), np.uint32)
    for f in range(4):
        pos[:4] = (corners[f % 2] + np.arange(4)[:, None] * 0.01).astype(np.float32)
        dirty = dev(torch, np.arange(4, dtype=np.int32)) if f else None
        bp.advance(dev(torch, pos), dq, dh, do, dirty=dirty)
        b = orc.box_bounds(pos, quat, half, off)
        want = orc.sap_pairs(b, body_id=body, dynamic=dyn)
        wf, wl = orc.found_lost(prev, want)
        if f:
            assert len(wf) > 8192 and len(wl) > 8192  # one key range over the LDS sort
        np.testing.assert_array_equal(bp.fetch(0), want, err_msg=f"pairs frame {f}")
        np.testing.assert_array_equal(bp.fetch(1), wf, err_msg=f"found frame {f}")
        np.testing.assert_array_equal(bp.fetch(2), wl, err_msg=f"lost frame {f}")
        prev = want
    np.testing.assert_equal(bp.counts()[3], 0)
    bp.close()


@pytest.mark.parametrize("n,seed,move", [(3000, 7, 20.0), (20000, 8, 1.0), (20000, 9, 40.0)])
def test_incremental_sap_matches_oracle(ctx, orc, torch_cuda, n, seed, move):
    _run_frames(ctx, orc, torch_cuda, n, 5, seed, 2.0, move, 0.1, True)


def test_incremental_sap_all_moving(ctx, orc, torch_cuda):
    # every box moves every frame (pile-like motion): moved-vs-moved crossings
    _run_frames(ctx, orc, torch_cuda, 8000, 4, 10, 1.5, 8.0, 1.0, False)


def test_sap_empty_and_single(ctx, torch_cuda):
    from fyzz_amd import sap
    torch = torch_cuda
    bp = sap.SapBroadphase(ctx, np.zeros(1, np.uint32), np.ones(1, np.uint8))
    one = dev(torch, np.zeros((1, 3), np.float32))
    bp.advance(one, dev(torch, np.array([[0, 0, 0, 1]], np.float32)), dev(torch, np.ones((1, 3), np.float32)),
               dev(torch, np.zeros(1, np.float32)))
    assert bp.counts()[:3] == (0, 0, 0)
    bp.close()
    bp0 = sap.SapBroadphase(ctx, np.zeros(0, np.uint32), np.zeros(0, np.uint8))
    e = dev(torch, np.zeros((0, 3), np.float32))
    bp0.advance(e, dev(torch, np.zeros((0, 4), np.float32)), e, dev(torch, np.zeros(0, np.float32)))
    assert bp0.counts()[:3] == (0, 0, 0)
    bp0.close()


def test_sap_million_boxes_properties(ctx, orc, torch_cuda):
    """1,000,000 AABBs (the broadphase-only workload at full size), 3 frames
    of 10% moves: endpoints sorted with unique payloads, and the pair list
    evolves exactly as (old - lost) + found, with every found pair
    overlapping and every lost pair not (checked on the host); and every
    frame's pairs, found and lost are bit-identical to the oracle's full
    sweep and set difference (so no newly overlapping pair goes
    unreported)."""
    from fyzz_amd import sap
    torch = torch_cuda
    n = 1_000_000
    half, pos = scenes.sap_scene(n)
    quat = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    off = np.full(n, 0.02, np.float32)
    bp = sap.SapBroadphase(ctx, np.arange(n, dtype=np.uint32), np.ones(n, np.uint8))
    dq, dh, do = dev(torch, quat), dev(torch, half), dev(torch, off)
    prev = None
    for f in range(3):
        dirty = None
        if f:
            slots, delta = scenes.sap_moves(n, f)
            pos[slots] = (pos[slots].astype(np.float64) + delta).astype(np.float32)
            dirty = dev(torch, slots.astype(np.int32))
        bp.advance(dev(torch, pos), dq, dh, do, dirty=dirty)
        pairs, found, lost = bp.fetch(0), bp.fetch(1), bp.fetch(2)
        vals, pays = bp.endpoints()
        v = u32(vals).astype(np.int64)
        p = u32(pays).astype(np.int64)
        key = v * (4 * n) + p
        assert np.all(np.diff(key, axis=1) > 0)
        assert np.all(np.sort(p, axis=1) == np.arange(2 * n))
        if prev is not None:
            sp, sf, sl = (set(map(tuple, x)) for x in (prev, found, lost))
            assert set(map(tuple, pairs)) == (sp - sl) | sf
            assert not (sf & sp) and sl <= sp
        # overlap of reported pairs on the exact keys
        b = u32(bp_bounds(ctx, torch, pos, quat, half, off))
        for arr, want in [(pairs, True), (found, True), (lost, False)]:
            if len(arr):
                a, c = arr[:, 0], arr[:, 1]
                ov = ~((b[0, a] > b[3, c]) | (b[0, c] > b[3, a]) | (b[1, a] > b[4, c]) | (b[1, c] > b[4, a]) |
                       (b[2, a] > b[5, c]) | (b[2, c] > b[5, a]))
                assert np.all(ov == want)
        # bit-exact against the oracle's sweep of the same poses
        want_pairs = orc.sap_pairs(orc.box_bounds(pos, quat, half, off))
        np.testing.assert_array_equal(pairs, want_pairs, err_msg=f"frame {f} pairs")
        if prev is not None:
            wf, wl = orc.found_lost(prev, want_pairs)
            np.testing.assert_array_equal(found, wf, err_msg=f"frame {f} found")
            np.testing.assert_array_equal(lost, wl, err_msg=f"frame {f} lost")
            print(f"frame {f}: {len(pairs)} pairs, {len(found)} found, {len(lost)} lost, bit-exact", flush=True)
        prev = pairs
    assert len(prev) > 1000
    bp.close()


def bp_bounds(ctx, torch, pos, quat, half, off):
    from fyzz_amd import sap
    b, _ = sap.box_bounds(ctx, dev(torch, pos), dev(torch, quat), dev(torch, half), dev(torch, off))
    torch.cuda.synchronize()
    return b


@pytest.mark.parametrize("n", [30000, 65537])
def test_incremental_sap_large_change_lists(ctx, orc, torch_cuda, n):
    # every box moves far: the per-frame candidate list exceeds the
    # one-workgroup path and the multi-block sort/unique path runs. The list
    # sort packs pair keys to 2B + 1 bits (slots < 2^B): 30000 boxes sort 32
    # bits in 4 passes, 65537 boxes 40 bits in 5 (an odd pass count, the
    # result coming back from the second buffer).
    _run_frames(ctx, orc, torch_cuda, n, 3, 12, 1.5, 15.0, 1.0, False, check_endpoints=False,
                delta_capacity=4_000_000 * max(1, n // 30000))
