"""GPU: the 100,000-box pile at full size (the throughput workload), spot
parity with the oracle while it falls, lands and settles.

The device world steps the seeded drop (graph replays between the spot
steps). At each spot step the device's synchronized state (body records,
live pair list with its manifold caches, plane caches, friction history)
is handed to the batched step oracle and the device takes one direct step;
world_parity.compare_step holds the bars (pairs, found/lost and partitions
bit-exact, contacts 1e-5, impulses and body state 1e-4). The last spot is
the formed pile (about 2e5 touching managers, 20+ partitions, overflow
waves and long crossing walks)."""
import pytest

from fyzz_amd import scenes
from oracle_bind import StepOracle
from world_parity import compare_step

pytestmark = pytest.mark.gpu


def test_pile_100k_spot_steps_match_oracle(ctx, orc, torch_cuda):
    from fyzz_amd.world import World
    n, dt = 100_000, 1 / 120
    sc = scenes.pile_scene(n)
    w = World(ctx, n + 1)
    try:
        fid = w.add_static_floor(0.0)
        w.set_body_material(fid, 0.5, 0.5, 0.0)
        w.add_dynamic_boxes(sc.half, sc.pos, sc.quat, friction=0.5, restitution=0.0)
        so = StepOracle(orc, sc.half, threads=16)
        spots = (1, 150, 400, 600)
        last = None
        for s in range(spots[-1] + 1):
            if s in spots:
                last = compare_step(w, so, dt, f"pile step {s}")
            else:
                w.step_graph(dt)
        managers, partitions, _ = last
        assert managers > 150_000 and partitions > 16
    finally:
        w.close()


def test_pile_100k_native_order_steps(ctx, torch_cuda):
    """The formed pile in the native solver order at full size: a world
    loaded with the state of a settled pile steps 8 times with the host
    island authority in the loop. Every step checks on its own that the
    device island pass and the authority sleep the same bodies, and that
    every manager the authority orders is touching; the solver list must
    hold exactly the touching managers of awake islands (the device order's
    selection of the same step, as a set)."""
    import numpy as np
    from fyzz_amd.world import World
    n, dt = 100_000, 1 / 120
    sc = scenes.pile_scene(n)
    w = World(ctx, n + 1)
    try:
        fid = w.add_static_floor(0.0)
        w.set_body_material(fid, 0.5, 0.5, 0.0)
        w.add_dynamic_boxes(sc.half, sc.pos, sc.quat, friction=0.5, restitution=0.0)
        for _ in range(600):
            w.step_graph(dt)
        recs = w.body_records()
    finally:
        w.close()

    def loaded(native):
        v = World(ctx, n + 1)
        f = v.add_static_floor(0.0)
        v.set_body_material(f, 0.5, 0.5, 0.0)
        v.add_dynamic_boxes(sc.half, recs["position"], recs["orientation"], linear=recs["linear_velocity"],
                            angular=recs["angular_velocity"], friction=0.5, restitution=0.0)
        v.set_solver_order(native)
        return v

    nat, dev = loaded(True), loaded(False)
    try:
        nat.step(dt)
        dev.step(dt)
        a, _, _ = nat.managers()
        b, _, _ = dev.managers()
        key = lambda m: set(zip(m["body0"].tolist(), m["body1"].tolist()))
        assert len(a) == len(b) and key(a) == key(b), "first step: the same touching managers in another order"
        assert not np.array_equal(a["body0"], b["body0"]), "the native order differs from the pair order"
        for _ in range(7):
            nat.step(dt)
        st = nat.stats()
        assert st["flags"] == 0 and st["managers"] > 150_000, st
    finally:
        nat.close()
        dev.close()
