"""GPU: island sleep and wake on the device world (the island pass of
world.hip: union-find islands, wakes through broadphase-connected
islands, sleep of islands whose bodies are all ready) against the
reference's own run (tests/golden/sleep_scene.npz; the oracle reproduces
it bit for bit) and, at every transition step, the whole step against the
oracle started from the device's state.

Bars: sleeping flags exact every step; states within 1e-4 relative (1e-6
absolute) of the reference's; contact events (notifications on) exact in
bodies, kinds, counts and shapes, impulse sums 1e-4; transition steps as
world_parity holds them."""
import numpy as np
import pytest

from author_cases import apply_edits
from test_author_gpu import _Ids, _world
from fyzz_amd import scenes
from oracle_bind import StepOracle
from world_parity import compare_step

from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden" / "sleep_scene.npz"
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [False, True])
def test_sleep_and_wake_match_reference(ctx, orc, torch_cuda, graph):
    g = np.load(GOLDEN)
    sc = scenes.Scene(g["half"], g["pos"], g["quat"], friction=0.5, restitution=0.0)
    steps = int(g["steps"][0])
    transitions = {49, 50, 51, 90, 91, 92, 100, 101, 110, 111}
    w = _world(ctx, sc)
    w.set_contact_notifications(True)
    so = StepOracle(orc, sc.half, threads=4)
    try:
        for s in range(steps):
            apply_edits(_Ids(w), g["edits"], s, to_id=lambda i: i + 1)
            rec = w.body_records()
            np.testing.assert_array_equal((rec["flags"] & 1).astype(np.uint8), g["sleeping"][s],
                                          err_msg=f"sleeping, step {s}")
            np.testing.assert_allclose(w.body_states(), g["state"][s], rtol=1e-4, atol=1e-6, err_msg=f"step {s}")
            if s in transitions and not graph:
                compare_step(w, so, 1 / 120, f"sleep step {s}")
            elif graph:
                w.step_graph(1 / 120)
            else:
                w.step(1 / 120)
            # events: persisting touches of sleeping islands are not reported
            ev, _ = w.contact_events()
            got = np.stack([ev["body0"], ev["body1"], ev["kind"], ev["contact_count"], ev["shape0"], ev["shape1"]],
                           1).astype(np.uint32)
            np.testing.assert_array_equal(got.reshape(-1, 6), g[f"ev_{s}"].reshape(-1, 6), err_msg=f"step {s} events")
            np.testing.assert_allclose(ev["normal_impulse"], g[f"evimp_{s}"], rtol=1e-4, atol=1e-5,
                                       err_msg=f"step {s} event impulses")
        np.testing.assert_allclose(w.body_states(), g["state"][steps], rtol=1e-4, atol=1e-6)
        assert w.stats()["flags"] == 0
    finally:
        w.close()
