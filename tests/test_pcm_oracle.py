"""CPU: the PCM oracle against the reference's golden vectors (candidates,
batch selection, GJK/EPA witnesses, and a stepped scene's manifolds) and
against closed forms (a box resting flat on the plane)."""
from pathlib import Path

import numpy as np
import pytest

from oracle_bind import OraclePcm, replay_pcm_scene

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def pcm(orc):
    return OraclePcm(orc.lib)


def test_box_box_candidates_match_reference(pcm):
    g = np.load(GOLDEN / "pcm_stateless.npz")
    cd = float(g["contact_distance"][0])
    many = 0
    for i in range(len(g["a"])):
        ok, c = pcm.candidates(g["a"][i], g["b"][i], cd)
        assert ok == bool(g["overlapping"][i]), i
        assert len(c) == g["count"][i], i
        np.testing.assert_array_equal(c, g["cands"][i][: len(c)], err_msg=f"pair {i}")
        if len(c) > 4:
            many += 1
            np.testing.assert_array_equal(pcm.batch_selection(c), g["sel"][i], err_msg=f"pair {i}")
    assert many > 10


def test_gjk_epa_matches_reference(pcm):
    g = np.load(GOLDEN / "pcm_stateless.npz")
    cd = float(g["contact_distance"][0])
    hits = 0
    for i in range(len(g["a"])):
        ok, c = pcm.gjk_epa(g["a"][i], g["b"][i], cd)
        assert ok == bool(g["gjk_ok"][i]), i
        if ok:
            hits += 1
            np.testing.assert_array_equal(c, g["gjk"][i], err_msg=f"pair {i}")
    assert hits > 100


def test_resting_box_plane_closed_form(pcm):
    # axis-aligned box of half extents h resting with penetration d: four
    # bottom corners at separation -d, positions (+-hx, 0, +-hz) + center
    h = np.array([0.5, 0.25, 0.75], np.float32)
    d = np.float32(0.01)
    box = np.concatenate([h, [1.0, h[1] - d, -2.0], [0, 0, 0, 1]]).astype(np.float32)
    cache = pcm.new_cache()
    n, pts, flags = pcm.plane_box(np.zeros(3, np.float32), box, 0.04, cache)
    assert n == 4 and flags & 1
    np.testing.assert_allclose(pts[:, 6], -d, rtol=1e-5)
    np.testing.assert_array_equal(pts[:, 3:6], np.tile([0, -1, 0], (4, 1)))
    got = sorted(map(tuple, np.round(pts[:, [0, 2]], 5)))
    want = sorted((1.0 + sx * 0.5, -2.0 + sz * 0.75) for sx in (-1, 1) for sz in (-1, 1))
    np.testing.assert_allclose(np.array(got), np.array(want), atol=1e-6)
    # unchanged pose: the cache is reused, no regeneration
    n2, pts2, flags2 = pcm.plane_box(np.zeros(3, np.float32), box, 0.04, cache)
    assert n2 == 4 and not (flags2 & 1)
    np.testing.assert_array_equal(pts2, pts)


def test_stacked_boxes_face_contact_closed_form(pcm):
    # identical axis-aligned boxes, one above the other with overlap d:
    # four contacts at separation -d, normal +y (from the lower, second
    # box toward the upper, first box)
    h = np.array([0.5, 0.5, 0.5], np.float32)
    d = np.float32(0.02)
    upper = np.concatenate([h, [0.0, 1.0 - d, 0.0], [0, 0, 0, 1]]).astype(np.float32)
    lower = np.concatenate([h, [0.0, 0.0, 0.0], [0, 0, 0, 1]]).astype(np.float32)
    n, pts, _ = pcm.box_box(upper, lower, 0.04, pcm.new_cache())
    assert n == 4
    np.testing.assert_allclose(pts[:, 6], -d, rtol=1e-4)
    np.testing.assert_allclose(pts[:, 3:6], np.tile([0, 1, 0], (4, 1)), atol=1e-6)


def test_scene_manifolds_match_reference(orc):
    g = np.load(GOLDEN / "pcm_scene.npz")
    checked = []

    def on_step(s, pairs, manifolds):
        if f"pairs_{s}" not in g.files:
            return
        np.testing.assert_array_equal(pairs, g[f"pairs_{s}"], err_msg=f"step {s} pairs")
        mp, mc, mpts = g[f"mpairs_{s}"], g[f"mcount_{s}"], g[f"mpoints_{s}"]
        assert len(mp) == len(manifolds), f"step {s}: {len(mp)} vs {len(manifolds)} manifolds"
        for k in range(len(mp)):
            key = (int(mp[k, 0]), int(mp[k, 1]))
            cnt, pts = manifolds[key]
            assert cnt == mc[k], f"step {s} pair {key}"
            np.testing.assert_array_equal(pts, mpts[k][:cnt], err_msg=f"step {s} pair {key}")
        checked.append(s)

    replay_pcm_scene(orc, g, on_step)
    assert len(checked) >= 15
