"""CPU: the contact-event oracle (Scene::publish_contact_events and the
touch transitions of update_contacts) against the reference's published
events for the golden pile, replayed from synchronized states with the
reference's descriptor order: found / persists / lost kinds, body and shape
order, point windows, outward normals and summed impulses, bit for bit."""
from pathlib import Path

import numpy as np

from oracle_bind import DyScene, OracleDy, oracle_contact_events

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_events_match_reference(orc):
    g = np.load(GOLDEN / "dy_scene.npz")
    e = np.load(GOLDEN / "dy_events.npz")
    sc = DyScene(orc, g)
    dy = OracleDy(orc.lib)
    history = {}
    acc_lin = np.zeros((sc.n, 3), np.float32)
    acc_ang = np.zeros((sc.n, 3), np.float32)
    prev = []
    kinds = np.zeros(3, np.int64)
    for s in range(sc.steps):
        mg, hist = sc.step_managers(s, history)
        out, patches, imp, _ = dy.step(1 / 120, sc.bodies(s, acc_lin, acc_ang), sc.statics, mg, hist)
        history = {k: patches[i] for i, k in enumerate(sc.keys[s])}
        acc_lin, acc_ang = out["sleep_linear_acc"].copy(), out["sleep_angular_acc"].copy()
        # every manager here is Gu-reversed: the floor (shape 0) is static,
        # and box/box pairs put the higher body id first
        ev, off, eimp, pt, pimp = oracle_contact_events(orc.lib, mg, imp, sc.keys[s], [True] * len(mg), prev)
        np.testing.assert_array_equal(ev, e[f"ev_{s}"], err_msg=f"step {s} events")
        np.testing.assert_array_equal(off, e[f"evoff_{s}"], err_msg=f"step {s} offsets")
        np.testing.assert_array_equal(eimp, e[f"evimp_{s}"], err_msg=f"step {s} impulses")
        np.testing.assert_array_equal(pt, e[f"pt_{s}"], err_msg=f"step {s} points")
        np.testing.assert_array_equal(pimp, e[f"ptimp_{s}"], err_msg=f"step {s} point impulses")
        kinds += np.bincount(ev[:, 2], minlength=3)
        prev = sorted(sc.keys[s])
    assert kinds[0] > 0 and kinds[1] > 0 and kinds[2] > 0, kinds
