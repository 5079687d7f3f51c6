"""CPU: the Dy oracle (pre-integration, partitions, contact prep with
correlated friction, PGS, integration, sleep check) against the compiled
reference's per-step body states, wake counters and conflict partitions,
replayed from synchronized start states; and closed forms (free fall,
resting box)."""
from pathlib import Path

import numpy as np

from oracle_bind import (DY_MANAGER, DY_STATIC, OracleDy, body_state_from, dy_bodies, replay_dy_scene)

GOLDEN = Path(__file__).resolve().parent / "golden"


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-6)


def test_dy_matches_reference_per_step(orc):
    g = np.load(GOLDEN / "dy_scene.npz")
    worst = {"state": 0.0}
    exact = [0, 0]

    def on_step(s, out, parts, mg):
        want = g["state"][s + 1]
        got = body_state_from(out)
        e = rel_err(got, want)
        worst["state"] = max(worst["state"], float(e.max()))
        exact[0] += int(np.array_equal(got, want))
        exact[1] += 1
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-6, err_msg=f"step {s}")
        np.testing.assert_array_equal(out["wake_counter"], g["wake"][s + 1], err_msg=f"step {s} wake")
        # partitions: the reference numbers them globally per task; local
        # = global - the owning task's first partition
        glob = g[f"part_{s}"].astype(np.int64)
        ends = g[f"task_{s}"].astype(np.int64)
        base = np.concatenate([[0], ends[:-1]])
        task = np.searchsorted(ends, glob, side="right")
        np.testing.assert_array_equal(parts.astype(np.int64), glob - base[task], err_msg=f"step {s} partitions")

    replay_dy_scene(orc, g, on_step)
    print(f"steps bit-identical: {exact[0]}/{exact[1]}, worst relative state error {worst['state']:.3g}")


def test_friction_writeback_matches_reference(orc):
    """The friction write-back (Scene::solved_friction_patches: anchor
    counts, broken flags, world anchor positions, f64 world impulses) of
    every step of the reference pile, bit for bit, matched by shape pair."""
    g = np.load(GOLDEN / "dy_scene.npz")
    seen = [0, 0]

    def on_step(s, out, parts, mg, an, keys):
        fr, fpos, fimp = g[f"fr_{s}"], g[f"frpos_{s}"], g[f"frimp_{s}"]
        want = {(int(a), int(b)): i for i, (a, b, _, _) in enumerate(fr)}
        assert len(want) == len(fr) == len(keys), f"step {s}: one patch per manager"
        for m, key in enumerate(keys):
            i = want[key]
            na = int(fr[i, 2])
            assert an["anchor_count"][m] == na and an["broken"][m] == fr[i, 3], f"step {s} {key}"
            np.testing.assert_array_equal(an["position"][m][:na], fpos[i][:na], err_msg=f"step {s} {key} positions")
            np.testing.assert_array_equal(an["impulse"][m][:na], fimp[i][:na], err_msg=f"step {s} {key} impulses")
            seen[0] += na
            seen[1] += int(np.any(fimp[i][:na] != 0.0))
    replay_dy_scene(orc, g, on_step, anchors=True)
    assert seen[0] > 1000 and seen[1] > 500, seen


def test_fp64_twin_close_to_fp32(orc):
    g = np.load(GOLDEN / "dy_scene.npz")
    errs = []

    def on_step(s, out, state64, mg):
        errs.append(float(rel_err(body_state_from(out), state64).max()))

    replay_dy_scene(orc, g, on_step, fp64=True)
    # the FP64 twin runs the same steps from the same FP32 start states
    assert max(errs) < 1e-2
    print("max FP32-vs-FP64 relative state error per step:", max(errs))


def test_free_fall_closed_form(orc):
    dy = OracleDy(orc.lib)
    b = dy_bodies(1)
    b["position"] = [0, 10, 0]
    b["inverse_mass"] = 1.0
    b["inverse_inertia"] = [1, 1, 1]
    b["angular_velocity"] = [0, 0, 0]
    dt = np.float32(1 / 120)
    st = np.zeros(1, DY_STATIC)
    out, _, _, _ = dy.step(float(dt), b, st, np.zeros(0, DY_MANAGER), np.zeros(0))
    vy = np.float32(0) + np.float32(-9.81) * dt
    assert out["linear_velocity"][0, 1] == vy
    assert out["position"][0, 1] == np.float32(10) + vy * dt


def _resting_box(points):
    b = dy_bodies(1)
    b["position"] = [0, 0.5, 0]
    b["inverse_mass"] = 1.0 / 1000.0
    b["inverse_inertia"] = [6.0 / 1000.0] * 3
    b["restitution"] = 0.0
    st = np.zeros(1, DY_STATIC)
    st["orientation"][0, 3] = 1
    st["static_friction"] = st["dynamic_friction"] = 0.5
    mg = np.zeros(1, DY_MANAGER)
    mg[0]["body0"], mg[0]["body1"], mg[0]["count"], mg[0]["history"] = 0, -1, len(points), 0xFFFFFFFF
    for i, (x, z) in enumerate(points):
        mg[0]["contacts"][i]["position"] = [x, 0, z]
        mg[0]["contacts"][i]["normal"] = [0, 1, 0]
        mg[0]["contacts"][i]["separation"] = 0.0
    return b, st, mg


def test_resting_box_single_contact_closed_form(orc):
    """One zero-separation contact under the center of mass: r x n = 0, so
    the row's response is 1/m and one Gauss-Seidel update cancels the
    gravity step exactly: impulse = m g dt, final velocity 0."""
    dy = OracleDy(orc.lib)
    b, st, mg = _resting_box([(0.0, 0.0)])
    out, patches, imp, _ = dy.step(1 / 120, b, st, mg, np.zeros(0))
    np.testing.assert_allclose(imp[0, 0], 1000.0 * np.float32(9.81) * np.float32(1 / 120), rtol=1e-6)
    np.testing.assert_allclose(out["linear_velocity"][0], 0.0, atol=1e-6)
    np.testing.assert_allclose(out["position"][0], [0, 0.5, 0], atol=1e-7)


def test_resting_box_four_contacts_converge(orc):
    """Four symmetric corner contacts: the projected Gauss-Seidel sweeps
    (4 position + 1 velocity) approach the balancing impulse m g dt from
    below; the residual fall speed is what the sweeps leave."""
    dy = OracleDy(orc.lib)
    b, st, mg = _resting_box([(0.5, 0.5), (0.5, -0.5), (-0.5, 0.5), (-0.5, -0.5)])
    out, patches, imp, _ = dy.step(1 / 120, b, st, mg, np.zeros(0))
    want = 1000.0 * 9.81 / 120
    assert want * 0.999 < imp.sum() <= want * 1.000001
    assert abs(out["linear_velocity"][0, 1]) < 1e-4
    assert patches[0]["anchor_count"] == 2
