"""The mid-run authoring scene of tests/golden/author_scene.npz: its
Scene record and the edits it applies between steps, replayed on any
world-like object with set_pose / set_velocity / set_material /
set_dynamics taking box slots or body ids. Test helper, not a test
module."""
from pathlib import Path

import numpy as np

from fyzz_amd import scenes

GOLDEN = Path(__file__).resolve().parent / "golden" / "author_scene.npz"


def load():
    g = np.load(GOLDEN)
    sc = scenes.Scene(g["half"], g["pos"], g["quat"], friction=0.5, restitution=0.0)
    return g, sc


def apply_edits(world, edits, step, to_id=lambda i: i):
    """Applies the golden's edits recorded before `step` (codes: 0 pose
    p3 q4, 1 velocity v3, 2 angular velocity v3, 3 material sf df rest,
    4 dynamics lin ang maxlin maxang gravity)."""
    for e in edits[edits[:, 0] == step]:
        code, i, v = int(e[1]), to_id(int(e[2])), e[3:]
        if code == 0:
            world.set_pose(i, v[0:3].astype(np.float32), v[3:7].astype(np.float32))
        elif code == 1:
            world.set_velocity(i, linear=v[0:3].astype(np.float32))
        elif code == 2:
            world.set_velocity(i, angular=v[0:3].astype(np.float32))
        elif code == 3:
            world.set_material(i, float(v[0]), float(v[1]), float(v[2]))
        elif code == 4:
            world.set_dynamics(i, float(v[0]), float(v[1]), float(v[2]), float(v[3]), bool(v[4]))
