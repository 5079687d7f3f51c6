"""CPU: authored edits between steps (Scene::set_pose, set_velocity,
set_angular_velocity, set_material, set_dynamics; reference
scene.cpp:589-781, 1022-1147) in the oracle's composed pipeline against
the reference's own run of the same scene and edits
(tests/golden/author_scene.npz). Each body of that scene takes part in at
most one dynamic/dynamic contact, so the manager order cannot change a
result and the pipeline must reproduce the reference bit for bit: the
cache reset and friction-history erase of a pose edit, the wakes, the
new material and damping all show in the states that follow."""
import numpy as np

from author_cases import apply_edits, load
from oracle_bind import OracleWorld


def test_authored_edits_match_reference_bit_exact(orc):
    g, sc = load()
    ow = OracleWorld(orc, sc)
    steps = int(g["steps"][0])
    state = g["state"]
    edits = g["edits"]
    for s in range(steps):
        apply_edits(ow, edits, s)
        np.testing.assert_array_equal(ow.body_states(), state[s], err_msg=f"start of step {s}")
        np.testing.assert_array_equal(ow.bodies["wake_counter"], g["wake"][s], err_msg=f"wake counters, step {s}")
        ow.step(1 / 120)
    np.testing.assert_array_equal(ow.body_states(), state[steps], err_msg="final state")
    # the edits moved things: a body lifted by the pose edit fell back
    assert state[30, 15, 1] > 0.85 and state[steps, 15, 1] < state[30, 15, 1] - 0.05


def test_pose_edit_resets_caches_and_history(orc):
    """Dropping the cache reset or the history erase of the small pose
    edits before step 25 changes the states that follow (the reference's
    states tell them apart)."""
    g, sc = load()
    edits = g["edits"]
    for skip in ("cache", "history"):
        ow = OracleWorld(orc, sc)
        for s in range(28):
            if s == 25:
                keep = (dict(ow.bb_caches), list(ow.pl_caches), dict(ow.history))
                apply_edits(ow, edits, s)
                if skip == "cache":
                    ow.bb_caches, ow.pl_caches = keep[0], keep[1]
                else:
                    ow.history = keep[2]
            else:
                apply_edits(ow, edits, s)
            ow.step(1 / 120)
        assert not np.array_equal(ow.body_states(), g["state"][28]), skip
