"""CPU: island sleep and wake in the oracle's composed pipeline against the
reference's own run (tests/golden/sleep_scene.npz): six two-box stacks
and four single boxes settle, every island goes to sleep at once
(Scene::deactivate_ready_accurate_islands, scene.cpp:6465-6536), then a
box dropped onto a sleeping one wakes it through the found pair and touch
(Scene::activate_published_islands, 6148-6262), an authored velocity
wakes a single box and an authored pose wakes a stack through its touching
peer. Each body stays in at most one dynamic/dynamic contact, so the
manager order cannot change a result: states, sleeping flags and wake
counters must equal the reference's bit for bit. The batched checker
(StepOracle, used by the GPU tests of whole piles) is held to the composed
pipeline through the same transitions."""
import numpy as np

from author_cases import apply_edits
from fyzz_amd import scenes
from oracle_bind import OracleWorld, StepOracle
from test_step_oracle import state_of

from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden" / "sleep_scene.npz"


def _scene(g):
    return scenes.Scene(g["half"], g["pos"], g["quat"], friction=0.5, restitution=0.0)


def test_sleep_and_wake_match_reference_bit_exact(orc):
    g = np.load(GOLDEN)
    ow = OracleWorld(orc, _scene(g))
    steps = int(g["steps"][0])
    for s in range(steps):
        apply_edits(ow, g["edits"], s)
        np.testing.assert_array_equal(ow.body_states(), g["state"][s], err_msg=f"state, step {s}")
        np.testing.assert_array_equal((ow.bodies["flags"] & 1).astype(np.uint8), g["sleeping"][s],
                                      err_msg=f"sleeping, step {s}")
        np.testing.assert_array_equal(ow.bodies["wake_counter"], g["wake"][s], err_msg=f"wake counters, step {s}")
        ow.step(1 / 120)
    np.testing.assert_array_equal(ow.body_states(), g["state"][steps])
    # the transitions happened: all asleep, then three wakes
    sl = g["sleeping"].sum(1).astype(np.int64)
    assert sl.max() == 16 and sl[-1] < 16 and (np.diff(sl) < 0).sum() >= 3


def test_events_through_sleep_match_reference(orc):
    """Contact events with notifications on: touches persisting inside a
    sleeping island are not reported (Scene::publish_contact_events,
    scene.cpp:9199-9215), found and lost touches are; bodies, kinds,
    counts and shapes bit for bit, impulse sums to the last bit of f64."""
    from oracle_bind import world_step_events
    g = np.load(GOLDEN)
    ow = OracleWorld(orc, _scene(g))
    prev = []
    quiet = 0
    for s in range(int(g["steps"][0])):
        apply_edits(ow, g["edits"], s)
        ow.step(1 / 120)
        ev, off, imp, pt, pimp, prev = world_step_events(orc.lib, ow, prev)
        want = g[f"ev_{s}"]
        np.testing.assert_array_equal(ev.reshape(-1, 6), want.reshape(-1, 6), err_msg=f"step {s} events")
        np.testing.assert_allclose(imp[:len(want)], g[f"evimp_{s}"], rtol=1e-12, err_msg=f"step {s} impulses")
        quiet += len(want) == 0
    assert quiet > 20  # steps whose islands all sleep report nothing


def test_islands_decide_sleep(orc):
    """Sleep is an island decision: with the stacks' islands ignored (each
    body deciding alone) the states differ from the reference's."""
    import oracle_bind
    g = np.load(GOLDEN)
    saved = oracle_bind._components
    try:
        oracle_bind._components = lambda n, edges: list(range(n))
        ow = OracleWorld(orc, _scene(g))
        differs = False
        for s in range(int(g["steps"][0])):
            apply_edits(ow, g["edits"], s)
            if not np.array_equal((ow.bodies["flags"] & 1).astype(np.uint8), g["sleeping"][s]):
                differs = True
                break
            ow.step(1 / 120)
        assert differs
    finally:
        oracle_bind._components = saved


def test_step_oracle_through_sleep_and_wake(orc):
    g = np.load(GOLDEN)
    sc = _scene(g)
    ow = OracleWorld(orc, sc)
    so = StepOracle(orc, sc.half, threads=4)
    checked = 0
    for s in range(int(g["steps"][0])):
        apply_edits(ow, g["edits"], s)
        if s == 0:
            ow.step(1 / 120)
            continue
        got = so.step(ow.bodies.copy(), ow.last[0].copy(), state_of(ow, s), 1 / 120)
        ow.step(1 / 120)
        assert got["managers"].tobytes() == ow.last[1].tobytes(), f"step {s} managers"
        assert got["bodies"].tobytes() == ow.bodies.tobytes(), f"step {s} bodies"
        checked += 1
    assert checked == int(g["steps"][0]) - 1
