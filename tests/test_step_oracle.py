"""CPU: the batched step oracle (StepOracle: the oracle's C++ batched PCM
entries, vectorized manager assembly and history lookup, one Dy call)
against the per-pair composed oracle pipeline (OracleWorld) on a small
pile. Started from the composed pipeline's own state after each step, one
StepOracle step must reproduce the pipeline's next step bit for bit:
pairs, found/lost, manifolds, manager list, partitions, impulses, bodies.
The GPU parity tests of the formed pile and the batched scenes use
StepOracle as their checker."""
import numpy as np

from fyzz_amd import scenes
from oracle_bind import DY_PATCH, OracleWorld, StepOracle


def small_pile(side=4, seed=scenes.SEED_BASE + 23):
    n = side ** 3
    idx = np.arange(n)
    i, k, j = idx % side, (idx // side) % side, idx // (side * side)
    half = np.stack([0.3 + 0.25 * scenes.uniform(seed, a, n) for a in range(3)], 1)
    pos = np.stack([(i - (side - 1) / 2) * 1.15, 0.62 + j * 1.15, (k - (side - 1) / 2) * 1.15], 1)
    quat = scenes.normalize_quat32(scenes.yaw_quat(scenes.uniform(seed, 3, n) * 2 * np.pi))
    vel = np.stack([scenes.normal(seed, 6 + a, n) * 0.3 for a in range(3)], 1)
    return scenes.Scene(half, pos, quat, vel=vel, friction=0.5, restitution=0.0)


def state_of(ow, step_no):
    """The composed pipeline's persistent state in the world's layout."""
    pairs = ow.last[0]
    n, p = ow.n, len(pairs)
    cbb = np.zeros((p, 56), np.float32)
    for i, (lo, hi) in enumerate(pairs):
        cbb[i] = ow.bb_caches[("bb", int(lo), int(hi))]
    hist = np.zeros(p + n, DY_PATCH)
    stamps = np.full(p + n, 0xFFFFFFFF, np.uint32)
    slot = {("bb", int(lo), int(hi)): i for i, (lo, hi) in enumerate(pairs)}
    for key, patch in ow.history.items():
        s = slot[key] if key[0] == "bb" else p + key[1]
        hist[s] = patch
        stamps[s] = step_no
    touch_bb = np.array([(int(lo), int(hi)) in ow.prev_touch_bb for lo, hi in pairs], np.uint32)
    touch_pl = np.array([i in ow.prev_touch_pl for i in range(n)], np.uint32)
    return {"caches_bb": cbb, "caches_pl": np.stack(ow.pl_caches).astype(np.float32), "history": hist,
            "stamps": stamps, "step": step_no, "touch_bb": touch_bb, "touch_pl": touch_pl}


def test_step_oracle_matches_composed_pipeline(orc):
    sc = small_pile()
    ow = OracleWorld(orc, sc)
    so = StepOracle(orc, sc.half, threads=4)
    dt = 1 / 120
    compared = 0
    for s in range(90):
        if s == 0:
            ow.step(dt)
            continue
        bodies = ow.bodies.copy()
        prev_pairs = ow.last[0].copy()
        state = state_of(ow, s)
        got = so.step(bodies, prev_pairs, state, dt)
        ow.step(dt)
        pairs, mg, imp, parts = ow.last
        np.testing.assert_array_equal(got["pairs"], pairs, err_msg=f"step {s} pairs")
        f, l = orc.found_lost(prev_pairs, pairs)
        np.testing.assert_array_equal(got["found"], f)
        np.testing.assert_array_equal(got["lost"], l)
        assert got["managers"].tobytes() == mg.tobytes(), f"step {s} managers"
        np.testing.assert_array_equal(got["partitions"], parts, err_msg=f"step {s} partitions")
        np.testing.assert_array_equal(got["impulses"], imp, err_msg=f"step {s} impulses")
        assert got["bodies"].tobytes() == ow.bodies.tobytes(), f"step {s} bodies"
        compared += len(mg)
    assert compared > 500


def test_step_oracle_scene_pairs_equal_filtered_sweep(orc):
    """With scenes, the oracle sweeps each scene on its own; the result is
    the whole sweep's pair set without the cross-scene pairs."""
    sc = small_pile(5)
    first = np.array([0, 40, 41, 90, sc.n])
    so = StepOracle(orc, sc.half, scene_first=first)
    b = np.zeros(sc.n, [("position", "<f4", 3), ("orientation", "<f4", 4)])
    b["position"], b["orientation"] = sc.pos, sc.quat
    got = so.pairs(b)
    allp = orc.sap_pairs(orc.box_bounds(sc.pos, sc.quat, sc.half, np.full(sc.n, 0.02, np.float32)))
    scene = np.searchsorted(first, np.arange(sc.n), side="right") - 1
    want = allp[scene[allp[:, 0]] == scene[allp[:, 1]]]
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    assert len(want) < len(allp)
    np.testing.assert_array_equal(got, want)
