"""GPU: the device step pipeline (fyzz_hip_world_*) against the oracle's
composed pipeline (oracle SAP -> PCM -> Dy in the world's manager order).

Each step starts both from the device's body records (synchronized state);
the oracle keeps its own manifold caches and friction history, which
therefore see the same inputs as the device's. Bars: broadphase pair
counts and the manager list (bodies, contact counts) exact; contacts 1e-5
relative (1e-6 absolute); partitions exact; body state after the step 1e-4
relative (1e-6 absolute). The quaternion update evaluates sin/cos, where
the reference's glibc sinf/cosf and the device differ in the last bit for
a fraction of arguments; the test reports how many steps are bit-identical.
"""
import numpy as np
import pytest

from fyzz_amd import scenes
from oracle_bind import OracleWorld

pytestmark = pytest.mark.gpu


def _add_statics(w, sc):
    for h, pos, quat, is_sphere in getattr(sc, "statics", []):
        sid = w.add_static_sphere(float(h), pos) if is_sphere else w.add_static_box(h, pos)
        w.set_body_pose(sid, pos, quat)
        w.set_body_material(sid, sc.friction, sc.friction, sc.restitution)


def make_world(ctx, sc):
    from fyzz_amd.world import World
    w = World(ctx, sc.n + 1)
    if sc.floor:
        fid = w.add_static_floor(sc.floor_height)
        w.set_body_material(fid, sc.friction, sc.friction, sc.restitution)
    if getattr(sc, "statics_first", False):
        _add_statics(w, sc)
    for i in range(sc.n):
        bid = (w.add_dynamic_sphere(float(sc.half[i, 0]), sc.density) if sc.sphere[i]
               else w.add_dynamic_box(sc.half[i], sc.density))
        w.set_body_pose(bid, sc.pos[i], sc.quat[i])
        w.set_body_velocity(bid, sc.vel[i])
        w.set_body_angular_velocity(bid, sc.ang[i])
        w.set_body_material(bid, sc.friction, sc.friction, sc.restitution)
    if not getattr(sc, "statics_first", False):
        _add_statics(w, sc)
    return w


def small_pile(side=3, seed=scenes.SEED_BASE + 23):
    n = side ** 3
    idx = np.arange(n)
    i, k, j = idx % side, (idx // side) % side, idx // (side * side)
    half = np.stack([0.3 + 0.25 * scenes.uniform(seed, a, n) for a in range(3)], 1)
    pos = np.stack([(i - (side - 1) / 2) * 1.15, 0.62 + j * 1.15, (k - (side - 1) / 2) * 1.15], 1)
    quat = scenes.normalize_quat32(scenes.yaw_quat(scenes.uniform(seed, 3, n) * 2 * np.pi))
    vel = np.stack([scenes.normal(seed, 6 + a, n) * 0.3 for a in range(3)], 1)
    return scenes.Scene(half, pos, quat, vel=vel, friction=0.5, restitution=0.0)


def compare_events(orc, w, ow, prev, s):
    """Device events == the event oracle over the oracle world's managers
    and impulses; returns this step's touching shape pairs."""
    from oracle_bind import world_step_events
    ev_o, off_o, eimp_o, pt_o, pimp_o, touching = world_step_events(orc.lib, ow, prev)
    ev, pt = w.contact_events()
    got = np.stack([ev["body0"], ev["body1"], ev["kind"], ev["contact_count"], ev["shape0"], ev["shape1"]], 1)
    np.testing.assert_array_equal(got.astype(np.uint32), ev_o, err_msg=f"step {s} events")
    np.testing.assert_array_equal(ev["point_offset"], off_o, err_msg=f"step {s} point offsets")
    np.testing.assert_allclose(ev["normal_impulse"], eimp_o, rtol=1e-4, atol=1e-5, err_msg=f"step {s} impulses")
    got_pt = np.concatenate([pt["position"], pt["normal"], pt["separation"][:, None]], 1)
    np.testing.assert_allclose(got_pt, pt_o, rtol=1e-5, atol=1e-6, err_msg=f"step {s} points")
    np.testing.assert_allclose(pt["normal_impulse"], pimp_o, rtol=1e-4, atol=1e-5, err_msg=f"step {s} point impulses")
    return touching, np.bincount(ev_o[:, 2], minlength=3)


def run_compare(ctx, orc, sc, steps, dt=1 / 120, events=False):
    w = make_world(ctx, sc)
    if events:
        w.set_contact_notifications(True)
    ow = OracleWorld(orc, sc)
    prev_touch = []
    kinds = np.zeros(3, np.int64)
    authored = w.body_records()
    np.testing.assert_array_equal(authored.view(np.uint8), ow.bodies.view(np.uint8), err_msg="authored records")
    exact_steps = 0
    for s in range(steps):
        ow.bodies = w.body_records().copy()
        w.step(dt)
        ow.step(dt)
        st = w.stats()
        assert st["flags"] == 0, st
        pairs, mg_o, imp_o, parts_o = ow.last
        assert st["pairs"] == len(pairs), f"step {s}"
        mg, imp, parts = w.managers()
        assert len(mg) == len(mg_o), f"step {s}: {len(mg)} vs {len(mg_o)} managers"
        np.testing.assert_array_equal(mg["body0"], mg_o["body0"], err_msg=f"step {s}")
        np.testing.assert_array_equal(mg["body1"], mg_o["body1"], err_msg=f"step {s}")
        np.testing.assert_array_equal(mg["count"], mg_o["count"], err_msg=f"step {s}")
        for f in ("position", "normal", "separation"):
            np.testing.assert_allclose(mg["contacts"][f], mg_o["contacts"][f], rtol=1e-5, atol=1e-6,
                                       err_msg=f"step {s} contact {f}")
        np.testing.assert_array_equal(parts, parts_o, err_msg=f"step {s} partitions")
        np.testing.assert_allclose(imp, imp_o, rtol=1e-4, atol=1e-5, err_msg=f"step {s} impulses")
        gs, os_ = w.body_states(), ow.body_states()
        np.testing.assert_allclose(gs, os_, rtol=1e-4, atol=1e-6, err_msg=f"step {s} bodies")
        exact_steps += int(np.array_equal(gs, os_))
        if events:
            prev_touch, k = compare_events(orc, w, ow, prev_touch, s)
            kinds += k
    if events:
        print("event kinds (found, persists, lost):", kinds.tolist())
        assert kinds.min() > 0
    return exact_steps


def test_small_pile_world_matches_oracle_pipeline(ctx, orc, torch_cuda):
    exact = run_compare(ctx, orc, small_pile(), 120, events=True)
    print(f"bit-identical steps: {exact}/120")


def test_grid_1000_world_matches_oracle_pipeline(ctx, orc, torch_cuda):
    """1,000 boxes on the 10x10x10 grid with seeded yaw and velocity over a
    floor: 240 steps at 1/120 s (the boxes fall, land and stack)."""
    exact = run_compare(ctx, orc, scenes.grid_scene(), 240, events=True)
    print(f"bit-identical steps: {exact}/240")


def test_plain_launch_matches_cooperative(torch_cuda):
    """The Dy grid-barrier kernels launched as plain kernels (the launch
    mode for tools that do not intercept cooperative launches) step a world
    bit-identically to cooperative launches: 40 steps of a 6x6x6 pile, body
    records compared after every step."""
    from fyzz_amd import Context
    sc = small_pile(6)
    ca, cb = Context(0), Context(0)
    cb.set_cooperative(False)
    wa, wb = make_world(ca, sc), make_world(cb, sc)
    try:
        for s in range(40):
            wa.step(1 / 120)
            wb.step(1 / 120)
            ra, rb = wa.body_records(), wb.body_records()
            assert ra.tobytes() == rb.tobytes(), f"step {s}"
        assert wb.stats()["flags"] == 0
        assert wb.stats()["managers"] > 0
    finally:
        wa.close()
        wb.close()
        ca.close()
        cb.close()


def test_graph_steps_match_eager_steps_sap_world(torch_cuda):
    """A world with the incremental SAP stepped through captured graphs (the
    first step eager, then two alternating graphs, with eager steps mixed in
    mid-run) against eager steps of an identical world: body states and
    world stats bit-identical after every step, eager work and a
    synchronization between replays."""
    from fyzz_amd import Context
    sc = small_pile(6)
    ca, cb = Context(0), Context(0)
    wa, wb = make_world(ca, sc), make_world(cb, sc)
    try:
        for s in range(40):
            wa.step(1 / 120)
            if 20 <= s < 24:
                wb.step(1 / 120)
            else:
                wb.step_graph(1 / 120)
            np.testing.assert_array_equal(wb.body_states(), wa.body_states(), err_msg=f"step {s}")
            assert wb.stats() == wa.stats(), f"step {s}"
        assert wa.stats()["managers"] > 0 and wa.stats()["flags"] == 0
    finally:
        wa.close()
        wb.close()
        ca.close()
        cb.close()


def _free_boxes(seed=scenes.SEED_BASE + 41, side=20, spacing=3.0):
    """side x side boxes (half extents U[0.2, 0.6] m) on a grid 3 m apart,
    0.6-2.0 m above the floor, with seeded yaw and small velocities: they
    land, tumble and settle without touching each other, so every step's
    managers share no body and the solve order cannot matter."""
    n = side * side
    idx = np.arange(n)
    half = np.stack([0.2 + 0.4 * scenes.uniform(seed, a, n) for a in range(3)], 1)
    pos = np.stack([(idx % side - (side - 1) / 2) * spacing, 0.6 + 1.4 * scenes.uniform(seed, 3, n) + half[:, 1],
                    (idx // side - (side - 1) / 2) * spacing], 1)
    yaw = scenes.uniform(seed, 4, n) * 2 * np.pi
    quat = scenes.normalize_quat32(scenes.yaw_quat(yaw))
    vel = np.stack([scenes.normal(seed, 5 + a, n) * 0.2 for a in range(3)], 1)
    ang = np.stack([scenes.normal(seed, 8 + a, n) * 0.5 for a in range(3)], 1)
    return scenes.Scene(half, pos, quat, vel, ang)


def test_world_matches_live_reference_when_order_cannot_matter(ctx, torch_cuda):
    """World::step end to end against the compiled NativeCpu scene, both
    running freely from the same authored start: 400 boxes that land,
    tumble and go to sleep without touching each other (each step's
    managers share no body, so the island order cannot change the solve).
    Every step: the same pair and manager counts, the same boxes asleep,
    body states within 1e-4 relative (1e-5 absolute); the steps that are
    bit-identical are reported."""
    import ref_live
    if not ref_live.available():
        pytest.skip("compiled reference (oracle/_ref) not built")
    sc = _free_boxes()
    w = make_world(ctx, sc)
    run = ref_live.RefRun(sc)
    same = 0
    worst = 0.0
    try:
        for s in range(300):
            w.step(1 / 120)
            st = run.step(1 / 120)
            ref, _wake, _ready, sleeping = st["next"]
            counts = w.stats()
            assert counts["flags"] == 0 and counts["pairs"] == len(st["pairs"]), f"step {s}: {counts}"
            assert counts["managers"] == len(st["managers"]), f"step {s}: {counts}"
            recs = w.body_records()
            got = np.concatenate([recs["position"], recs["linear_velocity"], recs["orientation"],
                                  recs["angular_velocity"]], 1)
            np.testing.assert_array_equal((recs["flags"] & 1).astype(bool), sleeping, err_msg=f"step {s} sleep")
            np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5, err_msg=f"step {s} states")
            worst = max(worst, float(np.abs(got.astype(np.float64) - ref).max()))
            same += int(np.array_equal(got, ref))
        assert int(sleeping.sum()) > sc.n // 2, "the boxes should be asleep by the end"
    finally:
        run.close()
        w.close()
    print(f"steps bit-identical to NativeCpu: {same}/300; largest difference {worst:.3g}; asleep at the end "
          f"{int(sleeping.sum())}/{sc.n}")


def _native_run(ctx, sc, steps, events=False):
    """The device world in the native solver order stepped freely beside the
    live reference from the same authored scene. Every step: the solver
    list equals the reference's descriptor stream, the same boxes sleep,
    body states bit-identical (checked against the 1e-4 relative / 1e-5
    absolute bar first, so a failure says how far apart they are); with
    events, the published contact events and points (f64 impulses
    included) equal the reference's too. Returns (steps whose states are
    bit-identical, most managers)."""
    import ref_live
    w = make_world(ctx, sc)
    w.set_solver_order(True)
    run = ref_live.RefRun(sc)
    if events:
        w.set_contact_notifications(True)
        run.set_contact_notifications(True)
    same, most = 0, 0
    kinds = np.zeros(3, np.int64)
    try:
        for s in range(steps):
            w.step(1 / 120)
            st = run.step(1 / 120)
            if events:
                ev, pt = w.contact_events()
                rev, roff, rimp, rpt, rpimp = run.contact_events()
                got = np.stack([ev["body0"], ev["body1"], ev["kind"], ev["contact_count"], ev["shape0"],
                                ev["shape1"]], 1).astype(np.uint32)
                np.testing.assert_array_equal(got, rev, err_msg=f"step {s} events")
                np.testing.assert_array_equal(ev["point_offset"], roff, err_msg=f"step {s} event point offsets")
                np.testing.assert_array_equal(ev["normal_impulse"], rimp, err_msg=f"step {s} event impulses")
                gpt = np.concatenate([pt["position"], pt["normal"], pt["separation"][:, None]], 1)
                np.testing.assert_array_equal(gpt, rpt, err_msg=f"step {s} event points")
                np.testing.assert_array_equal(pt["normal_impulse"], rpimp, err_msg=f"step {s} point impulses")
                kinds += np.bincount(rev[:, 2].astype(np.int64), minlength=3)
            mg, _, _ = w.managers()
            # ids: the floor 0, then the statics and the dynamic bodies in
            # creation order; a manager names static k as -2 - k, the floor -1
            ns = len(getattr(sc, "statics", []))
            first_static = getattr(sc, "statics_first", False)
            dyn0 = 1 + (ns if first_static else 0)
            st0 = 1 if first_static else 1 + sc.n
            b0 = mg["body0"].astype(np.int64) + dyn0
            b1 = mg["body1"].astype(np.int64)
            b1 = np.where(b1 == -1, 0, np.where(b1 < -1, st0 + (-2 - b1), b1 + dyn0))
            got = np.stack([np.minimum(b0, b1), np.maximum(b0, b1)], 1)
            want = np.asarray(st["keys"], np.int64).reshape(-1, 2)
            np.testing.assert_array_equal(got, want, err_msg=f"step {s}: solver order")
            ref, _wake, _ready, sleeping = st["next"]
            recs = w.body_records()
            np.testing.assert_array_equal((recs["flags"] & 1).astype(bool), sleeping, err_msg=f"step {s} sleep")
            states = w.body_states()
            np.testing.assert_allclose(states, ref, rtol=1e-4, atol=1e-5, err_msg=f"step {s} states")
            np.testing.assert_array_equal(states, ref, err_msg=f"step {s} states (bit-identical)")
            same += 1
            most = max(most, len(mg))
    finally:
        run.close()
        w.close()
    if events:
        assert kinds.min() > 0, f"found / persists / lost events all seen: {kinds.tolist()}"
    return same, most


@pytest.mark.parametrize("case", ["grid_1000", "hub", "sleep_wake", "pile_2000", "spheres_and_boxes", "statics",
                                  "statics_first"])
def test_native_order_world_matches_live_reference(ctx, torch_cuda, case):
    """World::step in the native solver order (the host island authority
    feeding the device solver list) against the live NativeCpu scene,
    end to end and free running, bit for bit: the 1,000-box grid over 240
    steps, the 65-partition hub over 60, the stack woken by dropped boxes
    over 600 (sleep, wakes through broadphase islands, re-sleep) and a
    2,000-box pile over 300 (one island of ~900 managers in 10+ tasks) and
    108 spheres with 108 boxes over 240 (sphere/sphere, sphere/box and
    sphere/plane contacts beside the box ones); on the grid, the sleep/wake
    stack and the spheres the contact events too."""
    import ref_live
    if not ref_live.available():
        pytest.skip("compiled reference (oracle/_ref) not built")
    sc, steps = {"grid_1000": (scenes.grid_scene(), 240), "hub": (ref_live.hub_scene(), 60),
                 "sleep_wake": (ref_live.wake_scene(), 600),
                 "pile_2000": (scenes.pile_scene(2000, footprint=15.0, height=160.0), 300),
                 "spheres_and_boxes": (ref_live.mixed_scene(), 240),
                 "statics": (ref_live.statics_scene(), 300),
                 "statics_first": (ref_live.statics_scene(first=True), 200)}[case]
    same, most = _native_run(ctx, sc, steps,
                             events=case in ("grid_1000", "sleep_wake", "spheres_and_boxes", "statics",
                                             "statics_first"))
    print(f"{case}: steps bit-identical to NativeCpu {same}/{steps}, most managers {most}")
    assert most > 40


def test_native_order_refuses_graphs_and_edits(ctx, torch_cuda):
    from fyzz_amd import FyzzError
    sc = small_pile()
    w = make_world(ctx, sc)
    w.set_solver_order(True)
    w.step(1 / 120)
    with pytest.raises(FyzzError) as e:
        w.step_graph(1 / 120)
    assert e.value.code == -4
    with pytest.raises(FyzzError):
        w.set_body_velocity(1, np.zeros(3, np.float32))
    with pytest.raises(FyzzError):
        w.set_solver_order(False)  # after the first step
    w.step(1 / 120)
    w.close()


def test_sphere_contact_cases_match_reference(ctx, torch_cuda):
    """First-step manifolds of the sphere contact cases (embedded centers
    and their exit-face ties, rotated boxes, edge contact, coincident
    centers, a gap below the contact distance, floor contacts, both Gu
    orders) against the live reference's, field by field and bit for bit:
    pairs, counts, contact positions, normals and separations."""
    import ref_live
    if not ref_live.available():
        pytest.skip("compiled reference (oracle/_ref) not built")
    sc = ref_live.sphere_cases()
    w = make_world(ctx, sc)
    w.set_solver_order(True)
    run = ref_live.RefRun(sc)
    try:
        w.step(1 / 120)
        st = run.step(1 / 120)
        mg, _, _ = w.managers()
        want = st["managers"]
        assert len(mg) == len(want) and len(mg) >= 16, (len(mg), len(want))
        for f in ("body0", "body1", "count"):
            np.testing.assert_array_equal(mg[f], want[f], err_msg=f)
        for f in ("position", "normal", "separation"):
            for k in range(4):
                live = mg["count"] > k
                np.testing.assert_array_equal(mg["contacts"][f][live, k], want["contacts"][f][live, k],
                                              err_msg=f"contact {k} {f}")
        np.testing.assert_array_equal(w.body_states(), st["next"][0])
    finally:
        run.close()
        w.close()
