"""CPU: the Dy oracle in the compiled reference's own solver order at scale,
stepped beside the live reference scene (oracle/_ref, built from the
reference's sources): the 1,000-box grid over 240 steps (boxes fall, land,
stack, a few go to sleep) and a hub slab carrying 64 box/box contacts (65
partitions: the overflow waves past the 32-bit colour mask) under a second
layer, 60 steps. Each step starts from the reference's synchronized state
after its island pass; the descriptor stream, its per-task partitions, the
next body states, wake counters and sleep flags must match bit for bit."""
import numpy as np
import pytest

import ref_live
from oracle_bind import OracleDy, body_state_from

pytestmark = pytest.mark.skipif(not ref_live.available(), reason="compiled reference (oracle/_ref) not built")


def _run(orc, scene, steps, min_parts):
    dy = OracleDy(orc.lib)
    seen = {"parts": 0, "mgrs": 0, "asleep": 0}

    def solve(s, st, b, statics, mg, hist):
        out, patches, imp, parts = dy.step(1 / 120, b, statics, mg, hist)
        np.testing.assert_array_equal(parts.astype(np.int64), st["partitions"], err_msg=f"step {s} partitions")
        nxt, wake, ready, sleeping = st["next"]
        np.testing.assert_array_equal(body_state_from(out), nxt, err_msg=f"step {s} states")
        np.testing.assert_array_equal(out["wake_counter"], wake, err_msg=f"step {s} wake counters")
        np.testing.assert_array_equal((out["flags"] & 1).astype(bool), sleeping, err_msg=f"step {s} sleep")
        np.testing.assert_array_equal((out["flags"] & 2).astype(bool) & ~sleeping, ready & ~sleeping,
                                      err_msg=f"step {s} ready")
        seen["parts"] = max(seen["parts"], int(parts.max()) + 1 if len(parts) else 0)
        seen["mgrs"] = max(seen["mgrs"], len(mg))
        seen["asleep"] = max(seen["asleep"], int(sleeping.sum()))
        return out, patches

    ref_live.replay(scene, steps, solve)
    assert seen["parts"] >= min_parts
    return seen


def test_grid_1000_in_reference_order(orc):
    from fyzz_amd import scenes
    seen = _run(orc, scenes.grid_scene(), 240, 11)
    assert seen["mgrs"] > 2000 and seen["asleep"] > 0
    print(seen)


def test_hub_overflow_waves_in_reference_order(orc):
    seen = _run(orc, ref_live.hub_scene(), 60, 65)
    assert seen["mgrs"] > 250
    print(seen)
