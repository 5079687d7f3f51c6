"""Spot-step parity of a device world against the batched step oracle
(StepOracle): the world's synchronized state goes to the oracle, the
device takes one direct step, and the results are compared at the bars
of the hot path:
  broadphase pairs (and found/lost)   bit-exact
  manager list (bodies, counts)       exact; contacts 1e-5 relative (1e-6 absolute)
  partitions                          exact
  contact impulses                    1e-4 relative (1e-5 absolute)
  friction write-back                 counts exact; anchor positions 1e-5,
                                      anchor impulses 1e-4 relative
  body state after the step           1e-4 relative (1e-6 absolute)
  sleeping and ready flags            exact
Test helper, not a test module."""
import time

import numpy as np

from oracle_bind import body_state_from


def compare_step(w, so, dt, label, reports=True):
    """One spot step; returns (managers, partitions, bit-identical bodies)."""
    bodies = w.body_records()
    prev_pairs = w.broadphase_reports(0)
    state = w.persistent_state()
    w.step(dt)
    t0 = time.perf_counter()
    want = so.step(bodies, prev_pairs, state, dt)
    t_orc = time.perf_counter() - t0
    st = w.stats()
    assert st["flags"] == 0, (label, st)
    for which, key in ((0, "pairs"), (1, "found"), (2, "lost")):
        if which and not reports:
            continue
        got = w.broadphase_reports(which)
        np.testing.assert_array_equal(got, want[key], err_msg=f"{label} {key}")
    mg, imp, parts = w.managers()
    wm = want["managers"]
    assert len(mg) == len(wm), f"{label}: {len(mg)} vs {len(wm)} managers"
    for f in ("body0", "body1", "count"):
        np.testing.assert_array_equal(mg[f], wm[f], err_msg=f"{label} manager {f}")
    # friction history: the device names the slot, the oracle a compacted index
    np.testing.assert_array_equal(mg["history"] != 0xFFFFFFFF, wm["history"] != 0xFFFFFFFF,
                                  err_msg=f"{label} manager history")
    for f in ("position", "normal", "separation"):
        np.testing.assert_allclose(mg["contacts"][f], wm["contacts"][f], rtol=1e-5, atol=1e-6,
                                   err_msg=f"{label} contact {f}")
    np.testing.assert_array_equal(parts, want["partitions"], err_msg=f"{label} partitions")
    np.testing.assert_allclose(imp, want["impulses"], rtol=1e-4, atol=1e-5, err_msg=f"{label} impulses")
    # friction write-back: anchor counts and broken flags exact, positions
    # at the contact bar, impulses at the impulse bar
    an, wa = w.friction_anchors(), want["anchors"]
    np.testing.assert_array_equal(an["anchor_count"], wa["anchor_count"], err_msg=f"{label} anchor counts")
    np.testing.assert_array_equal(an["broken"], wa["broken"], err_msg=f"{label} broken patches")
    live = np.arange(2)[None, :] < wa["anchor_count"][:, None]
    np.testing.assert_allclose(an["position"][live], wa["position"][live], rtol=1e-5, atol=1e-6,
                               err_msg=f"{label} anchor positions")
    np.testing.assert_allclose(an["impulse"][live], wa["impulse"][live], rtol=1e-4, atol=1e-5,
                               err_msg=f"{label} anchor impulses")
    rec = w.body_records()
    np.testing.assert_array_equal(rec["flags"] & 1, want["bodies"]["flags"] & 1, err_msg=f"{label} sleeping")
    np.testing.assert_array_equal(rec["flags"] & 2, want["bodies"]["flags"] & 2, err_msg=f"{label} ready")
    got_b, want_b = w.body_states(), body_state_from(want["bodies"])
    np.testing.assert_allclose(got_b, want_b, rtol=1e-4, atol=1e-6, err_msg=f"{label} bodies")
    exact = int(np.sum(np.all(got_b == want_b, axis=1)))
    nparts = int(parts.max()) + 1 if len(parts) else 0
    found = f" ({len(want['found'])} found, {len(want['lost'])} lost)" if reports else ""
    print(f"{label}: {len(want['pairs'])} pairs{found}, {len(wm)} managers, {nparts} partitions, "
          f"{exact}/{len(got_b)} bodies bit-identical; oracle {t_orc:.1f} s", flush=True)
    return len(wm), nparts, exact
