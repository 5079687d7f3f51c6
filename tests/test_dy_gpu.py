"""GPU: the HIP Dy stage (partitions, contact prep, PGS, integration, sleep
check) against the oracle and the reference goldens.

Bars: partitions bit-exact; body pose and velocities within 1e-4 relative
(1e-6 absolute floor for values near zero); contact impulses within 1e-4
relative (1e-6 absolute). The kernels keep the reference's FP32 operation
order and the quaternion update computes sin/cos as the host C library
does, so body states are bit-identical where the inputs are (asserted in
the reference-order replays; the other tests report the count).
"""
from pathlib import Path

import numpy as np
import pytest

from dy_cases import random_case, with_history
from oracle_bind import DyScene, OracleDy, body_state_from

GOLDEN = Path(__file__).resolve().parent / "golden"
pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-4, 1e-6


def _solver(ctx, nb, nm):
    from fyzz_amd.dy import DySolver
    return DySolver(ctx, nb, nm)


def _gpu_step(solver, dt, bodies, statics, mg, hist, anchors=False):
    from fyzz_amd.dy import DY_BODY, DY_PATCH, from_device, to_device
    db = to_device(bodies)
    ds = to_device(statics)
    dm = to_device(mg)
    dh = to_device(hist) if len(hist) else None
    patches, imp, parts = solver.step(dt, db, ds, dm, dh)
    flags, nparts, rounds = solver.status()
    out = (from_device(db, DY_BODY), from_device(patches, DY_PATCH), imp.cpu().numpy(),
           parts.cpu().numpy().astype(np.uint32), (flags, nparts, rounds))
    if anchors:
        out = out + (solver.friction_anchors(len(mg)),)
    return out


def _compare_anchors(got, want, label):
    """Friction write-back: counts and broken flags exact, positions at the
    contact bar (1e-5), impulses at the impulse bar (1e-4)."""
    np.testing.assert_array_equal(got["anchor_count"], want["anchor_count"], err_msg=f"{label} anchor counts")
    np.testing.assert_array_equal(got["broken"], want["broken"], err_msg=f"{label} broken")
    live = np.arange(2)[None, :] < want["anchor_count"][:, None]
    np.testing.assert_allclose(got["position"][live], want["position"][live], rtol=1e-5, atol=1e-6,
                               err_msg=f"{label} anchor positions")
    np.testing.assert_allclose(got["impulse"][live], want["impulse"][live], rtol=1e-4, atol=1e-5,
                               err_msg=f"{label} anchor impulses")


def test_partitions_match_reference(ctx, orc, torch_cuda):
    from fyzz_amd.dy import to_device
    g = np.load(GOLDEN / "dy_scene.npz")
    sc = DyScene(orc, g)
    solver = _solver(ctx, sc.n, 256)
    acc = np.zeros((sc.n, 3), np.float32)
    for s in range(sc.steps):
        mg, _ = sc.step_managers(s, {})
        parts = solver.partition(to_device(sc.bodies(s, acc, acc)), to_device(mg)).cpu().numpy()
        solver.status()
        glob = g[f"part_{s}"].astype(np.int64)
        ends = g[f"task_{s}"].astype(np.int64)
        base = np.concatenate([[0], ends[:-1]])
        task = np.searchsorted(ends, glob, side="right")
        np.testing.assert_array_equal(parts.astype(np.int64), glob - base[task], err_msg=f"step {s}")


def test_scene_steps_match_oracle_and_reference(ctx, orc, torch_cuda):
    """The golden pile's 72 steps from synchronized start states: GPU ==
    oracle on the same inputs (oracle history carried), GPU == reference
    next state; the friction write-back == the oracle's (which equals the
    reference's bit for bit)."""
    g = np.load(GOLDEN / "dy_scene.npz")
    sc = DyScene(orc, g)
    odd = OracleDy(orc.lib)
    solver = _solver(ctx, sc.n, 256)
    history = {}
    acc_lin = np.zeros((sc.n, 3), np.float32)
    acc_ang = np.zeros((sc.n, 3), np.float32)
    same = total = 0
    from fyzz_amd.dy import FRICTION_ANCHORS
    for s in range(sc.steps):
        mg, hist = sc.step_managers(s, history)
        bodies = sc.bodies(s, acc_lin, acc_ang)
        wa = np.zeros(max(len(mg), 1), FRICTION_ANCHORS)
        want, wp, wi, wparts = odd.step(1 / 120, bodies, sc.statics, mg, hist, anchors=wa)
        got, gp, gi, gparts, st, ga = _gpu_step(solver, 1 / 120, bodies, sc.statics, mg, hist, anchors=True)
        _compare_anchors(ga, wa[:len(mg)], f"step {s}")
        np.testing.assert_array_equal(gparts, wparts, err_msg=f"step {s} partitions")
        gs, ws = body_state_from(got), body_state_from(want)
        np.testing.assert_allclose(gs, ws, rtol=RTOL, atol=ATOL, err_msg=f"step {s} vs oracle")
        np.testing.assert_allclose(gs, g["state"][s + 1], rtol=RTOL, atol=ATOL, err_msg=f"step {s} vs reference")
        np.testing.assert_array_equal(got["flags"], want["flags"], err_msg=f"step {s} ready flags")
        np.testing.assert_allclose(got["wake_counter"], want["wake_counter"], rtol=RTOL, atol=ATOL)
        np.testing.assert_allclose(gi, wi, rtol=RTOL, atol=ATOL, err_msg=f"step {s} impulses")
        np.testing.assert_array_equal(gp["anchor_count"], wp["anchor_count"])
        np.testing.assert_array_equal(gp["flags"], wp["flags"])
        same += int((gs == ws).all(axis=1).sum())
        total += sc.n
        history = {k: wp[i] for i, k in enumerate(sc.keys[s])}
        acc_lin = want["sleep_linear_acc"].copy()
        acc_ang = want["sleep_angular_acc"].copy()
    print(f"bodies bit-identical to the oracle: {same}/{total}")


@pytest.mark.parametrize("seed", [1, 2])
def test_random_managers_two_steps_match_oracle(ctx, orc, torch_cuda, seed):
    """2,048 bodies, ~5,600 managers in random order including a 70-manager
    hub (overflow wave) and dynamic/static managers against two statics;
    two steps, the second correlating with the first step's patches."""
    bodies, statics, mg = random_case(seed)
    odd = OracleDy(orc.lib)
    solver = _solver(ctx, len(bodies), len(mg))
    hist = np.zeros(0)
    for step in range(2):
        want, wp, wi, wparts = odd.step(1 / 120, bodies, statics, mg, hist)
        got, gp, gi, gparts, (flags, nparts, rounds) = _gpu_step(solver, 1 / 120, bodies, statics, mg, hist)
        assert flags == 0
        assert nparts == int(wparts.max()) + 1 and nparts > 32
        np.testing.assert_array_equal(gparts, wparts, err_msg=f"step {step} partitions")
        gs, ws = body_state_from(got), body_state_from(want)
        np.testing.assert_allclose(gs, ws, rtol=RTOL, atol=ATOL, err_msg=f"step {step} bodies")
        np.testing.assert_allclose(gi, wi, rtol=RTOL, atol=1e-5, err_msg=f"step {step} impulses")
        np.testing.assert_array_equal(gp["anchor_count"], wp["anchor_count"])
        for f in ("body0_normal", "body1_normal", "body0_anchors", "body1_anchors"):
            np.testing.assert_allclose(gp[f], wp[f], rtol=RTOL, atol=ATOL, err_msg=f)
        print(f"step {step}: partitions {nparts}, rounds {rounds}, bodies bit-identical "
              f"{int((gs == ws).all(axis=1).sum())}/{len(gs)}")
        bodies = want.copy()
        mg, hist = with_history(mg, wp)


def test_free_fall_and_empty(ctx, orc, torch_cuda):
    """No managers: pre-integration and integration only (closed form)."""
    from fyzz_amd.dy import DY_MANAGER, DY_STATIC, default_bodies
    b = default_bodies(3)
    b["position"] = [[0, 10, 0], [1, 5, 0], [2, 7, 0]]
    b["inverse_mass"] = 1.0
    b["inverse_inertia"] = 1.0
    b["angular_damping"] = 0.0
    b["flags"][2] |= 1  # asleep: untouched
    solver = _solver(ctx, 8, 8)
    got, _, _, _, st = _gpu_step(solver, 1 / 120, b, np.zeros(1, DY_STATIC), np.zeros(0, DY_MANAGER), np.zeros(0))
    dt = np.float32(1 / 120)
    vy = np.float32(-9.81) * dt
    np.testing.assert_array_equal(got["linear_velocity"][:2, 1], [vy, vy])
    np.testing.assert_array_equal(got["position"][:2, 1], np.float32([10, 5]) + vy * dt)
    np.testing.assert_array_equal(got[2:].view(np.uint8), b[2:].view(np.uint8))


def test_malformed_manager_is_reported(ctx, orc, torch_cuda):
    from fyzz_amd import FyzzError
    from fyzz_amd.dy import DY_MANAGER, DY_STATIC, default_bodies, to_device
    b = default_bodies(4)
    mg = np.zeros(1, DY_MANAGER)
    mg[0]["body0"], mg[0]["body1"], mg[0]["count"] = 1, 9, 1  # body 9 does not exist
    solver = _solver(ctx, 8, 8)
    solver.step(1 / 120, to_device(b), to_device(np.zeros(1, DY_STATIC)), to_device(mg))
    with pytest.raises(FyzzError) as e:
        solver.status()
    assert e.value.code == -1


def test_overflow_waves_with_ten_thousand_pending(ctx, orc, torch_cuda):
    """48 hub bodies with 250 dynamic/dynamic managers each (built on the
    CPU): the first 32 of a hub fill its first partition mask and the
    other 218 wait for overflow waves (10,464 pending in all), whose
    pending list the device sorts in parallel before assigning waves of
    32. Partitions bit-exact against the oracle's sequential first-fit;
    the step's state within the bar. (One hub of 10k managers would need
    ~10k partitions, beyond the 4,096-partition capacity: that case is
    the capacity error.)"""
    bodies, statics, mg = random_case(7, nb=12_000, n_dd=1500, hub=250, n_static=300, hubs=48)
    want, wp, wi, wparts = OracleDy(orc.lib).step(1 / 120, bodies, statics, mg, np.zeros(0))
    solver = _solver(ctx, len(bodies), len(mg))
    got, gp, gi, gparts, (flags, nparts, rounds) = _gpu_step(solver, 1 / 120, bodies, statics, mg, np.zeros(0))
    assert flags == 0
    assert nparts == int(wparts.max()) + 1 and nparts > 6 * 32
    np.testing.assert_array_equal(gparts, wparts, err_msg="partitions")
    np.testing.assert_allclose(body_state_from(got), body_state_from(want), rtol=RTOL, atol=ATOL)
    print(f"partitions {nparts}, rounds {rounds}")


@pytest.mark.parametrize("case", ["grid_1000", "hub_overflow"])
def test_reference_order_at_scale(ctx, orc, torch_cuda, case):
    """The HIP Dy stage fed the live compiled reference's descriptor stream
    (its island order) from its synchronized post-island-pass states: the
    1,000-box grid over 240 steps and the 65-partition hub slab over 60.
    Partitions bit-exact against the reference's per-task numbering, next
    states, wake counters and impulses at the bars above against the
    reference (states) and the oracle on the same inputs (impulses), sleep
    and ready flags exact; the oracle's history is carried."""
    import ref_live
    if not ref_live.available():
        pytest.skip("compiled reference (oracle/_ref) not built")
    from fyzz_amd import scenes
    scene, steps = (scenes.grid_scene(), 240) if case == "grid_1000" else (ref_live.hub_scene(), 60)
    odd = OracleDy(orc.lib)
    solver = _solver(ctx, scene.n, 4096)
    count = {"same": 0, "total": 0, "parts": 0}

    def solve(s, st, b, statics, mg, hist):
        want, wp, wi, wparts = odd.step(1 / 120, b, statics, mg, hist)
        got, gp, gi, gparts, (flags, nparts, rounds) = _gpu_step(solver, 1 / 120, b, statics, mg, hist)
        assert flags == 0, f"step {s} status {flags}"
        np.testing.assert_array_equal(gparts.astype(np.int64), st["partitions"], err_msg=f"step {s} partitions")
        nxt, wake, ready, sleeping = st["next"]
        gs = body_state_from(got)
        np.testing.assert_allclose(gs, nxt, rtol=RTOL, atol=ATOL, err_msg=f"step {s} vs reference")
        np.testing.assert_allclose(got["wake_counter"], wake, rtol=RTOL, atol=ATOL, err_msg=f"step {s} wake")
        np.testing.assert_array_equal(got["flags"], want["flags"], err_msg=f"step {s} flags")
        np.testing.assert_array_equal((got["flags"] & 1).astype(bool), sleeping, err_msg=f"step {s} sleep")
        np.testing.assert_allclose(gi, wi, rtol=RTOL, atol=1e-5, err_msg=f"step {s} impulses")
        count["same"] += int((gs == nxt).all(axis=1).sum())
        count["total"] += scene.n
        count["parts"] = max(count["parts"], int(nparts))
        return want, wp

    ref_live.replay(scene, steps, solve)
    assert count["parts"] >= (65 if case == "hub_overflow" else 11)
    assert count["same"] == count["total"], "every body state bit-identical to the reference"
    print(f"{case}: bodies bit-identical to the reference {count['same']}/{count['total']}, "
          f"most partitions {count['parts']}")
