"""GPU: authored edits after the first step (fyzz_hip_world_set_body_pose,
_velocity, _material, _dynamics; reference Scene::set_pose, set_velocity,
set_angular_velocity, set_material, set_dynamics, scene.cpp:589-781,
1022-1147) on the device world, against the reference's own run of the
same scene and edits (tests/golden/author_scene.npz, which the oracle
reproduces bit for bit) and against the oracle's pipeline stepped from
the device's synchronized state right after each edit.

Bars: every step's state within 1e-4 relative (1e-6 absolute) of the
reference's; at the edit steps the device's whole next step (pairs,
managers, contacts, partitions, impulses, state) as world_parity holds
it against the oracle started from the device's post-edit state."""
import numpy as np
import pytest

from author_cases import apply_edits, load
from oracle_bind import StepOracle
from world_parity import compare_step

pytestmark = pytest.mark.gpu


class _Ids:
    """World calls by body id (floor 0, box slot i = id i + 1)."""

    def __init__(self, w):
        self.w = w

    def set_pose(self, i, p, q):
        self.w.set_body_pose(i, p, q)

    def set_velocity(self, i, linear=None, angular=None):
        if linear is not None:
            self.w.set_body_velocity(i, linear)
        if angular is not None:
            self.w.set_body_angular_velocity(i, angular)

    def set_material(self, i, sf, df, rest):
        self.w.set_body_material(i, sf, df, rest)

    def set_dynamics(self, i, lin, ang, maxlin, maxang, gravity):
        self.w.set_body_dynamics(i, linear_damping=lin, angular_damping=ang, max_linear_velocity=maxlin,
                                 max_angular_velocity=maxang, gravity_enabled=1 if gravity else 0)


def _world(ctx, sc):
    from fyzz_amd.world import World
    w = World(ctx, sc.n + 1)
    fid = w.add_static_floor(0.0)
    w.set_body_material(fid, 0.5, 0.5, 0.0)
    w.add_dynamic_boxes(sc.half, sc.pos, sc.quat, friction=0.5, restitution=0.0)
    return w


def test_authored_edits_match_reference(ctx, orc, torch_cuda):
    g, sc = load()
    edits, state = g["edits"], g["state"]
    steps = int(g["steps"][0])
    edit_steps = set(int(s) for s in edits[:, 0])
    w = _world(ctx, sc)
    so = StepOracle(orc, sc.half, threads=4)
    worst = 0.0
    try:
        for s in range(steps):
            apply_edits(_Ids(w), edits, s, to_id=lambda i: i + 1)
            got = w.body_states()
            np.testing.assert_allclose(got, state[s], rtol=1e-4, atol=1e-6, err_msg=f"start of step {s}")
            worst = max(worst, float(np.abs(got - state[s]).max()))
            if s in edit_steps:
                # the oracle from the device's post-edit records, caches and
                # history: the edit's invalidations must be on the device
                compare_step(w, so, 1 / 120, f"edit step {s}")
            else:
                w.step(1 / 120)
        np.testing.assert_allclose(w.body_states(), state[steps], rtol=1e-4, atol=1e-6, err_msg="final state")
        print(f"authored edits: {steps} steps within the bar of the reference (max abs diff {worst:.3g})")
    finally:
        w.close()


def test_authoring_errors(ctx, torch_cuda):
    """The reference's rejections, before and after the first step."""
    from fyzz_amd._lib import FyzzError
    g, sc = load()
    w = _world(ctx, sc)
    try:
        for stepped in (False, True):
            with pytest.raises(FyzzError) as e:
                w.set_body_pose(1, np.zeros(3, np.float32), np.array([0, 0, 0, 1.5], np.float32))
            assert e.value.code == -1  # not a sane pose
            with pytest.raises(FyzzError) as e:
                w.set_body_pose(0, np.zeros(3, np.float32), np.array([0, 0, 0, 1], np.float32))
            assert e.value.code == -4  # the floor does not move
            with pytest.raises(FyzzError) as e:
                w.set_body_material(1, 0.5, 0.5, 1.5)
            assert e.value.code == -1
            with pytest.raises(FyzzError) as e:
                w.set_body_dynamics(1, position_iterations=0)
            assert e.value.code == -1
            with pytest.raises(FyzzError) as e:
                w.set_body_dynamics(1, lock_flags=1)
            assert e.value.code == -4  # outside the admitted class
            if not stepped:
                w.step(1 / 120)
        assert w.stats()["flags"] == 0
    finally:
        w.close()
