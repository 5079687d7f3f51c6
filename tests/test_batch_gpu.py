"""GPU: the batched form (independent scenes in one device world, per-scene
broadphase, one launch per stage over all scenes) against worlds that hold
one scene each. Scenes never interact, the per-scene broadphase uses the
broadphase's own key encoding and pair filter, and the manager order keeps
each scene's relative order, so every scene must step bit-identically to
its own world (the bar is 1e-4 relative on velocity; the test asks for
equality)."""
import numpy as np
import pytest

from fyzz_amd import scenes

pytestmark = pytest.mark.gpu


def make(ctx, half, pos, quat, counts=None):
    from fyzz_amd.world import World
    w = World(ctx, len(half) + 1)
    fid = w.add_static_floor(0.0)
    w.set_body_material(fid, 0.5, 0.5, 0.0)
    w.add_dynamic_boxes(half, pos, quat, friction=0.5, restitution=0.0)
    if counts is not None:
        w.set_scenes(counts)
    return w


def test_batched_scenes_match_single_scene_worlds(ctx, torch_cuda):
    counts, half, pos, quat = scenes.batch_scenes(6, seed=scenes.SEED_BASE + 50, min_boxes=3, max_boxes=120,
                                                  height=3.0, footprint_per_100k=1500.0)
    first = np.concatenate([[0], np.cumsum(counts.astype(np.int64))]).astype(np.int64)
    batched = make(ctx, half, pos, quat, counts)
    singles = [make(ctx, half[a:b], pos[a:b], quat[a:b]) for a, b in zip(first[:-1], first[1:])]
    for s in range(90):
        batched.step(1 / 120)
        for w in singles:
            w.step(1 / 120)
        st = batched.stats()
        assert st["flags"] == 0, st
        single = [w.stats() for w in singles]
        assert st["pairs"] == sum(x["pairs"] for x in single), f"step {s}"
        assert st["managers"] == sum(x["managers"] for x in single), f"step {s}"
        got = batched.body_states()
        for k, w in enumerate(singles):
            np.testing.assert_array_equal(got[first[k]:first[k + 1]], w.body_states(),
                                          err_msg=f"step {s} scene {k} ({counts[k]} boxes)")
    assert batched.stats()["managers"] > 0
    print("scenes", counts.tolist(), "managers at the end", batched.stats()["managers"])


def test_scene_too_large_is_rejected(ctx, torch_cuda):
    from fyzz_amd import FyzzError
    counts, half, pos, quat = scenes.batch_scenes(1, min_boxes=1100, max_boxes=1100, height=50.0,
                                                  footprint_per_100k=3000.0)
    w = make(ctx, half, pos, quat, counts)
    with pytest.raises(FyzzError) as e:
        w.step(1 / 120)
    assert e.value.code == -1


def test_graph_steps_match_eager_steps(ctx, torch_cuda):
    """Graph-captured steps (two alternating graphs, replayed) against
    eager steps of an identical world, with eager work and synchronization
    between replays; every replay is checked."""
    counts, half, pos, quat = scenes.batch_scenes(5, seed=scenes.SEED_BASE + 51, min_boxes=3, max_boxes=80,
                                                  height=3.0, footprint_per_100k=1500.0)
    eager = make(ctx, half, pos, quat, counts)
    graphed = make(ctx, half, pos, quat, counts)
    for s in range(40):
        eager.step(1 / 120)
        graphed.step_graph(1 / 120)
        a, b = eager.body_states(), graphed.body_states()
        np.testing.assert_array_equal(b, a, err_msg=f"step {s}")
        assert graphed.stats() == eager.stats(), f"step {s}"
    assert eager.totals()["managers"] > 0


def test_autotune_keeps_results(ctx, torch_cuda):
    """Launch shapes picked by the autotuner do not change results: a tuned
    world and an untuned one stepped the same number of times agree."""
    counts, half, pos, quat = scenes.batch_scenes(4, seed=scenes.SEED_BASE + 52, min_boxes=10, max_boxes=60,
                                                  height=3.0, footprint_per_100k=1500.0)
    tuned = make(ctx, half, pos, quat, counts)
    plain = make(ctx, half, pos, quat, counts)
    pb, sb = tuned.autotune(1 / 120, trial_steps=2)
    assert pb > 0 and sb > 0
    steps = tuned.totals()["steps"]
    for _ in range(steps):
        plain.step(1 / 120)
    for _ in range(10):
        tuned.step_graph(1 / 120)
        plain.step(1 / 120)
    np.testing.assert_array_equal(tuned.body_states(), plain.body_states())


def test_set_workgroups_keeps_results(ctx, torch_cuda):
    """Launch shapes set by hand (clamped to residency; 0 keeps a count)
    do not change results: worlds at 16/32 and 64/512 partition/PGS
    workgroups step the same as the default shapes."""
    counts, half, pos, quat = scenes.batch_scenes(4, seed=scenes.SEED_BASE + 53, min_boxes=10, max_boxes=60,
                                                  height=3.0, footprint_per_100k=1500.0)
    ref = make(ctx, half, pos, quat, counts)
    small = make(ctx, half, pos, quat, counts)
    wide = make(ctx, half, pos, quat, counts)
    assert small.set_workgroups(16, 32) == (16, 32)
    pb, sb = wide.set_workgroups(64, 512)
    assert pb == 64 and 256 <= sb <= 512
    assert wide.set_workgroups(0, 0) == (pb, sb)
    with pytest.raises(Exception):
        wide.set_workgroups(-1, 0)
    for _ in range(40):
        ref.step(1 / 120)
        small.step_graph(1 / 120)
        wide.step(1 / 120)
    np.testing.assert_array_equal(small.body_states(), ref.body_states())
    np.testing.assert_array_equal(wide.body_states(), ref.body_states())


def test_batched_4096_scenes_spot_steps_match_oracle(ctx, orc, torch_cuda):
    """The batched workload at full size: 4,096 scenes of U{8..512} boxes
    (about 1.07M) in one world, graph replays between spot steps; at each
    spot the whole world's step is compared with the oracle pipeline run
    scene by scene from the device's synchronized state (pairs and
    partitions bit-exact, no cross-scene pair, contacts 1e-5, impulses and
    velocities 1e-4; world_parity.compare_step)."""
    from oracle_bind import StepOracle
    from world_parity import compare_step
    counts, half, pos, quat = scenes.batch_scenes(4096)
    first = np.concatenate([[0], np.cumsum(counts.astype(np.int64))]).astype(np.int64)
    w = make(ctx, half, pos, quat, counts)
    try:
        so = StepOracle(orc, half, scene_first=first, threads=16)
        spots = (1, 200)
        for s in range(spots[-1] + 1):
            if s in spots:
                managers, _, _ = compare_step(w, so, 1 / 120, f"4096 scenes, step {s}", reports=False)
            else:
                w.step_graph(1 / 120)
        assert managers > 100_000
    finally:
        w.close()


def test_batched_scenes_with_spheres_match_single_scene_worlds(ctx, torch_cuda):
    """Scenes holding spheres beside boxes: the batched world (per-scene
    broadphase with sphere bounds) steps every scene bit-identically to a
    world holding only that scene."""
    import ref_live
    from fyzz_amd.world import World
    parts = [ref_live.mixed_scene(side=s, seed=scenes.SEED_BASE + 70 + s) for s in (2, 3, 4, 3)]

    def author(w, sc, shift):
        for i in range(sc.n):
            bid = (w.add_dynamic_sphere(float(sc.half[i, 0]), sc.density) if sc.sphere[i]
                   else w.add_dynamic_box(sc.half[i], sc.density))
            w.set_body_pose(bid, sc.pos[i] + np.float32([shift, 0, 0]), sc.quat[i])
            w.set_body_velocity(bid, sc.vel[i])
            w.set_body_material(bid, sc.friction, sc.friction, sc.restitution)

    def world(scs, counts=None):
        w = World(ctx, sum(sc.n for sc in scs) + 1)
        fid = w.add_static_floor(0.0)
        w.set_body_material(fid, 0.5, 0.5, 0.0)
        for k, sc in enumerate(scs):
            author(w, sc, 0.0)  # scenes overlap in space; only the broadphase keeps them apart
        if counts is not None:
            w.set_scenes(counts)
        return w

    batched = world(parts, [sc.n for sc in parts])
    singles = [world([sc]) for sc in parts]
    first = np.concatenate([[0], np.cumsum([sc.n for sc in parts])])
    try:
        for s in range(120):
            batched.step(1 / 120)
            for w in singles:
                w.step(1 / 120)
            assert batched.stats()["flags"] == 0
            got = batched.body_states()
            for k, w in enumerate(singles):
                np.testing.assert_array_equal(got[first[k]:first[k + 1]], w.body_states(), err_msg=f"step {s} scene {k}")
        assert batched.stats()["managers"] > 0
    finally:
        batched.close()
        for w in singles:
            w.close()
