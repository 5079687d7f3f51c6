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
